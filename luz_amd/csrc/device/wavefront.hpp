// Wavefront integrator (stages over SoA path state). Placeholder until the staged
// kernels land: selecting it reports LUZGPU_E_UNSUPPORTED.
#pragma once

#include "integrator.hpp"
#include <stdexcept>

namespace luzgpu
{
	struct	WavefrontState
	{
		void	release() {}
	};

	inline void	renderWavefront(WavefrontState&, const SceneView&, const GpuParams&, float*, unsigned long long*, hipStream_t)
	{
		throw std::runtime_error("Unsupported: wavefront integrator not built yet");
	}
}
