// Scene-file front end: the subset of the reference's .luz / OBJ / Radiance-HDR readers
// that the device path consumes. Semantics follow the reference loader so the same file
// produces the same FP64 scene on both sides:
//   sections and blank-line termination  src/io/SceneFile/SceneFile.cpp:256-321
//   [settings] keys                       src/io/SceneFile/Settings.cpp:162-640
//   [spectra] + reflectance(...) colours  src/io/SceneFile/Parsing.cpp:136-174,292-368,
//                                         src/core/ColorScience.cpp:222-321
//   material blocks                       src/io/SceneFile/Materials.cpp:26-70,760-870,927-960
//   [meshes]                              src/io/SceneFile/Meshes.cpp:4-74
//   primitives / object blocks            src/io/SceneFile/Objects.cpp:357-406,800-880,1502-1620
//   sphere / area / sphere-light blocks   src/io/SceneFile/Objects.cpp:420-660,880-1160
//   volume blocks and volume presets      src/io/SceneFile/Objects.cpp:690-800,1224-1330,
//                                         src/core/MeasuredMaterials.cpp:139-152,225-330
//   OBJ faces (fan triangulation)         src/io/OBJReader.cpp:576-641
//   degenerate-triangle removal           src/io/OBJReader.cpp:418-445, Mesh.cpp:360-385
//   RGBE decode + sRGB->ACEScg            src/core/EnvironmentMap.cpp:140-150,376-454
// Unsupported keys raise an error naming the key (the reference is strict the same way).
#include "scene.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <map>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <utility>
#include <vector>

namespace luzgpu
{
	double	luminance(const dvec3& c)
	{
		// ColorManagement.cpp:313-320
		return (c[0] * 0.2722287168 + c[1] * 0.6740817658 + c[2] * 0.0536895174);
	}

	namespace
	{
		// Largest atmosphere march (view samples, and light samples per view sample) the device
		// path accepts: 4096 x 4096 FP64 steps is already seconds per miss-ray wave.
		constexpr int	kMaxAtmosphereSamples = 4096;

		// The CIE 1931 colour-matching fits ColorScience.cpp:19-42 uses (three / two / two
		// piecewise Gaussians).
		double	cmfGaussian(double v) { return std::exp(-0.5 * v * v); }
		double	cmfX(double w)
		{
			return 0.362 * cmfGaussian((w - 442.0) * (w < 442.0 ? 0.0624 : 0.0374))
				+ 1.056 * cmfGaussian((w - 599.8) * (w < 599.8 ? 0.0264 : 0.0323))
				- 0.065 * cmfGaussian((w - 501.1) * (w < 501.1 ? 0.0490 : 0.0382));
		}
		double	cmfY(double w)
		{
			return 0.821 * cmfGaussian((w - 568.8) * (w < 568.8 ? 0.0213 : 0.0247))
				+ 0.286 * cmfGaussian((w - 530.9) * (w < 530.9 ? 0.0613 : 0.0322));
		}
		double	cmfZ(double w)
		{
			return 1.217 * cmfGaussian((w - 437.0) * (w < 437.0 ? 0.0845 : 0.0278))
				+ 0.681 * cmfGaussian((w - 459.0) * (w < 459.0 ? 0.0385 : 0.0725));
		}

		// XYZ -> ACEScg (ColorManagement.cpp:139-143, 272-275)
		dvec3	acescgFromXYZ(double x, double y, double z)
		{
			return dvec3{1.6410233797 * x + -0.3248032942 * y + -0.2364246952 * z,
				-0.6636628587 * x + 1.6153315917 * y + 0.0167563477 * z,
				0.0117218943 * x + -0.0082844420 * y + 0.9883948585 * z};
		}

		// XYZ -> ACEScg, then non-negative channels scaled to a maximum of 1
		// (normalizeVisibleColor, ColorScience.cpp:49-60)
		dvec3	visibleFromXYZ(double x, double y, double z)
		{
			dvec3 c = acescgFromXYZ(x, y, z);
			for (double& v : c) v = std::isfinite(v) ? std::max(0.0, v) : 0.0;
			const double top = std::max(c[0], std::max(c[1], c[2]));
			if (top <= 0.0 || !std::isfinite(top))
			{
				throw std::runtime_error("Spectral color converted outside the visible RGB gamut.");
			}
			for (double& v : c) v /= top;
			return c;
		}
	}

	dvec3	blackbodyColor(double kelvin)
	{
		if (!std::isfinite(kelvin) || kelvin < 1000.0 || kelvin > 40000.0)
		{
			throw std::runtime_error("Blackbody temperature must be finite and within 1000-40000 K.");
		}
		// relativePlanckRadiance summed every 5 nm over 360-830 nm (ColorScience.cpp:77-88, 196-214)
		double x = 0.0, y = 0.0, z = 0.0;
		for (double w = 360.0; w <= 830.0; w += 5.0)
		{
			const double meters = w * 1.0e-9;
			const double exponent = 1.438776877e-2 / (meters * kelvin);
			const double r = exponent > 700.0 ? 0.0 : 1.0 / (std::pow(meters, 5.0) * std::expm1(exponent));
			x += r * cmfX(w);
			y += r * cmfY(w);
			z += r * cmfZ(w);
		}
		return visibleFromXYZ(x, y, z);
	}

	dvec3	wavelengthColor(double nanometers)
	{
		if (!std::isfinite(nanometers) || nanometers < 360.0 || nanometers > 830.0)
		{
			throw std::runtime_error("Wavelength must be finite and within 360-830 nm.");
		}
		return visibleFromXYZ(cmfX(nanometers), cmfY(nanometers), cmfZ(nanometers));
	}

	namespace
	{
		std::string	trim(const std::string& s)
		{
			std::size_t a = 0, b = s.size();
			while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) a++;
			while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) b--;
			return (s.substr(a, b - a));
		}

		std::string	lower(std::string s)
		{
			for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
			return (s);
		}

		double	parseDouble(const std::string& text, const std::string& what)
		{
			const std::string t = trim(text);
			char* end = nullptr;
			const double v = std::strtod(t.c_str(), &end);
			if (t.empty() || end != t.c_str() + t.size() || !std::isfinite(v))
			{
				throw std::runtime_error("Invalid " + what + ": " + text);
			}
			return (v);
		}

		// "(x,y,z)" with strtod semantics (the reference reads these with sscanf %lf).
		dvec3	parseTriple(const std::string& text, const std::string& what)
		{
			std::string t = trim(text);
			std::string lt = lower(t);
			if (lt.rfind("acescg(", 0) == 0)
			{
				t = t.substr(6);
			}
			if (t.size() < 2 || t.front() != '(' || t.back() != ')')
			{
				throw std::runtime_error("Unsupported " + what + " value (only (x,y,z), acescg(...), srgb(...) and linear_srgb(...) are supported): " + text);
			}
			const std::string body = t.substr(1, t.size() - 2);
			dvec3 out{};
			std::size_t start = 0;
			for (int i = 0; i < 3; i++)
			{
				const std::size_t comma = body.find(',', start);
				if ((i < 2) != (comma != std::string::npos))
				{
					throw std::runtime_error("Invalid " + what + ": " + text);
				}
				out[i] = parseDouble(body.substr(start, comma == std::string::npos ? std::string::npos : comma - start), what);
				start = comma + 1;
			}
			return (out);
		}

		// ColorManagement::acescgFromLinearSRGB (ColorManagement.cpp:222-225): the linear-sRGB
		// to ACEScg matrix (:121-125) after clamping to non-negative finite values
		dvec3	acescgFromLinearSRGB(const dvec3& c)
		{
			static const double M[3][3] = {
				{0.6131324224, 0.3395380158, 0.0474166960},
				{0.0701243808, 0.9163940113, 0.0134515239},
				{0.0205876575, 0.1095745716, 0.8697854040}
			};
			auto nonNegative = [](double v) { return (std::isfinite(v) && v > 0.0) ? v : 0.0; };
			const double r = nonNegative(c[0]), g = nonNegative(c[1]), b = nonNegative(c[2]);
			dvec3 o;
			for (int i = 0; i < 3; i++) o[i] = M[i][0] * r + M[i][1] * g + M[i][2] * b;
			return (o);
		}

		// ColorManagement::srgbToLinear (ColorManagement.cpp:186-201)
		double	srgbToLinear(double v)
		{
			if (!std::isfinite(v) || v <= 0.0) return (0.0);
			if (v >= 1.0) return (1.0);
			if (v <= 0.04045) return (v / 12.92);
			return (std::pow((v + 0.055) / 1.055, 2.4));
		}

		// ColorScience::reflectanceCurve (ColorScience.cpp:222-267): the samples (nm, value),
		// sorted and linearly interpolated, held flat past either end, are weighted every 5 nm
		// over 360-830 nm by a 6504 K Planck illuminant and the CIE fits; the ACEScg colour of
		// that sum is divided channel by channel by the illuminant's own, each ratio clamped
		// to [0,1].
		dvec3	reflectanceColor(std::vector<std::pair<double, double>> samples)
		{
			if (samples.size() < 2)
			{
				throw std::invalid_argument("Spectral reflectance curves require at least two samples.");
			}
			for (const auto& s : samples)
			{
				if (!std::isfinite(s.first) || s.first < 360.0 || s.first > 830.0)
				{
					throw std::invalid_argument("Spectral reflectance wavelengths must be finite and within 360-830 nm.");
				}
				if (!std::isfinite(s.second) || s.second < 0.0 || s.second > 1.0)
				{
					throw std::invalid_argument("Spectral reflectance values must be finite and in [0,1].");
				}
			}
			std::sort(samples.begin(), samples.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
			for (std::size_t i = 1; i < samples.size(); i++)
			{
				if (samples[i].first <= samples[i - 1].first)
				{
					throw std::invalid_argument("Spectral reflectance wavelengths must be unique.");
				}
			}
			auto at = [&](double w) -> double
			{
				if (w <= samples.front().first) return samples.front().second;
				if (w >= samples.back().first) return samples.back().second;
				std::size_t hi = 1;
				while (samples[hi].first <= w) hi++;	// first sample strictly above w
				const auto& a = samples[hi - 1];
				const auto& b = samples[hi];
				const double span = b.first - a.first;
				const double t = span > 0.0 ? (w - a.first) / span : 0.0;
				return (a.second * (1.0 - t) + b.second * t);
			};
			double x = 0.0, y = 0.0, z = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
			for (double w = 360.0; w <= 830.0; w += 5.0)
			{
				const double meters = w * 1.0e-9;
				const double exponent = 1.438776877e-2 / (meters * 6504.0);
				const double light = exponent > 700.0 ? 0.0 : 1.0 / (std::pow(meters, 5.0) * std::expm1(exponent));
				const double r = at(w);
				const double ex = light * cmfX(w), ey = light * cmfY(w), ez = light * cmfZ(w);
				x += r * ex;
				y += r * ey;
				z += r * ez;
				wx += ex;
				wy += ey;
				wz += ez;
			}
			const dvec3 c = acescgFromXYZ(x, y, z);
			const dvec3 white = acescgFromXYZ(wx, wy, wz);
			dvec3 out{};
			for (int i = 0; i < 3; i++)
			{
				if (!(white[i] > 0.0) || !std::isfinite(white[i]))
				{
					throw std::invalid_argument("Spectral reflectance illuminant produced an invalid white point.");
				}
				const double v = c[i] / white[i];
				out[i] = (!std::isfinite(v) || v <= 0.0) ? 0.0 : (v >= 1.0 ? 1.0 : v);
			}
			return (out);
		}

		// One "nm value" line of a spectrum ('#' starts a comment, ',' and ';' read as
		// spaces; Parsing.cpp:322-366, ColorScience.cpp:279-312). False for an empty line;
		// throws when the line is not exactly two numbers.
		bool	spectrumSample(std::string line, std::pair<double, double>& sample, const std::string& where)
		{
			const std::size_t hash = line.find('#');
			if (hash != std::string::npos) line.erase(hash);
			std::replace(line.begin(), line.end(), ',', ' ');
			std::replace(line.begin(), line.end(), ';', ' ');
			line = trim(line);
			if (line.empty()) return (false);
			std::istringstream in(line);
			std::string trailing;
			if (!(in >> sample.first >> sample.second) || (in >> trailing))
			{
				throw std::runtime_error("Invalid spectral reflectance sample in " + where + ". Use wavelength_nm,value.");
			}
			return (true);
		}

		// ColorScience::loadReflectanceCurve (ColorScience.cpp:269-321): a text file of samples.
		dvec3	reflectanceFileColor(const std::string& path)
		{
			std::ifstream in(path);
			if (!in)
			{
				throw std::runtime_error("Spectral reflectance curve could not be opened: " + path);
			}
			std::vector<std::pair<double, double>> samples;
			std::string line;
			std::size_t number = 0;
			while (std::getline(in, line))
			{
				number++;
				std::pair<double, double> s;
				if (spectrumSample(line, s, path + ":" + std::to_string(number))) samples.push_back(s);
			}
			try
			{
				return (reflectanceColor(samples));
			}
			catch (const std::exception& e)
			{
				throw std::runtime_error("Invalid spectral reflectance curve '" + path + "': " + e.what());
			}
		}

		// What a colour value may refer to beyond its own text: the [spectra] section read so
		// far and the scene file's folder (Parsing.cpp:161-174). Set while a file is read.
		struct	ColorContext
		{
			std::map<std::string, dvec3>	spectra;
			std::filesystem::path	base;
		};
		thread_local const ColorContext*	tColorContext = nullptr;

		std::string	resolvePath(const std::filesystem::path& base, const std::string& file);

		// reflectance / reflectance_curve / spectrum / spectral(NAME or PATH), the name or path
		// optionally quoted: a [spectra] entry of that name, else a curve file beside the scene
		// (Parsing.cpp:136-174).
		dvec3	spectralColor(const std::string& t, const std::string& function, const std::string& what)
		{
			const std::string prefix = function + "(";
			if (t.back() != ')')
			{
				throw std::runtime_error("Invalid " + what + " color function.");
			}
			std::string a = trim(t.substr(prefix.size(), t.size() - prefix.size() - 1));
			if (a.size() >= 2 && ((a.front() == '"' && a.back() == '"') || (a.front() == '\'' && a.back() == '\'')))
			{
				a = a.substr(1, a.size() - 2);
			}
			if (a.empty())
			{
				throw std::runtime_error("Invalid " + what + " spectral curve reference.");
			}
			if (tColorContext)
			{
				const auto it = tColorContext->spectra.find(a);
				if (it != tColorContext->spectra.end()) return (it->second);
			}
			return (reflectanceFileColor(resolvePath(tColorContext ? tColorContext->base : std::filesystem::path(), a)));
		}

		// A colour value (Parsing.cpp:227-264): (r,g,b) and acescg(r,g,b) are ACEScg as given;
		// srgb / display_srgb / display(r,g,b) are sRGB-encoded and linear_srgb / linearsrgb /
		// lin_srgb(r,g,b) linear sRGB, both converted to ACEScg.
		dvec3	parseColor(const std::string& text, const std::string& what)
		{
			const std::string t = trim(text);
			const std::string lt = lower(t);
			for (const char* name : {"reflectance", "reflectance_curve", "spectrum", "spectral"})
			{
				if (lt.rfind(std::string(name) + "(", 0) == 0) return (spectralColor(t, name, what));
			}
			for (const char* name : {"srgb", "display_srgb", "display"})
			{
				const std::string prefix = std::string(name) + "(";
				if (lt.rfind(prefix, 0) == 0)
				{
					const dvec3 e = parseTriple(t.substr(prefix.size() - 1), what);
					return (acescgFromLinearSRGB(dvec3{srgbToLinear(e[0]), srgbToLinear(e[1]), srgbToLinear(e[2])}));
				}
			}
			for (const char* name : {"linear_srgb", "linearsrgb", "lin_srgb"})
			{
				const std::string prefix = std::string(name) + "(";
				if (lt.rfind(prefix, 0) == 0)
				{
					return (acescgFromLinearSRGB(parseTriple(t.substr(prefix.size() - 1), what)));
				}
			}
			// single-argument colour functions with an optional unit suffix (Parsing.cpp:108-135, 191-210)
			auto argument = [&](const std::string& name, const std::string& suffix) -> double
			{
				const std::string prefix = name + "(";
				if (lt.back() != ')') throw std::runtime_error("Invalid " + what + " function.");
				std::string a = trim(t.substr(prefix.size(), t.size() - prefix.size() - 1));
				const std::string la = lower(a);
				if (la.size() > suffix.size() && la.compare(la.size() - suffix.size(), suffix.size(), suffix) == 0)
				{
					a = trim(a.substr(0, a.size() - suffix.size()));
				}
				std::size_t used = 0;
				double v = 0.0;
				try
				{
					v = std::stod(a, &used);
				}
				catch (const std::exception&)
				{
					throw std::runtime_error("Invalid " + what + " function value.");
				}
				if (used != a.size() || !std::isfinite(v)) throw std::runtime_error("Invalid " + what + " function value.");
				return v;
			};
			if (lt.rfind("wavelength(", 0) == 0) return (wavelengthColor(argument("wavelength", "nm")));
			for (const char* name : {"blackbody", "temperature", "color_temperature"})
			{
				if (lt.rfind(std::string(name) + "(", 0) == 0) return (blackbodyColor(argument(name, "k")));
			}
			if (lt == "solar" || lt == "sun") return (blackbodyColor(5778.0));
			return (parseTriple(t, what));
		}

		bool	splitAssignment(const std::string& line, std::string& key, std::string& value)
		{
			const std::size_t eq = line.find('=');
			if (eq == std::string::npos)
			{
				return (false);
			}
			key = lower(trim(line.substr(0, eq)));
			value = trim(line.substr(eq + 1));
			return (true);
		}

		bool	namedBlock(const std::string& line, const std::string& keyword, std::string& name)
		{
			const std::string t = trim(line);
			if (t.rfind(keyword + " ", 0) != 0 || t.back() != '{')
			{
				return (false);
			}
			// _parseNamedBlockHeader (Parsing.cpp:42-65): one word, no blanks or braces
			name = trim(t.substr(keyword.size(), t.size() - keyword.size() - 1));
			return (!name.empty() && name.find_first_of(" \t{}") == std::string::npos);
		}

		std::string	resolvePath(const std::filesystem::path& base, const std::string& file)
		{
			std::filesystem::path p(file);
			if (p.is_absolute())
			{
				return (p.string());
			}
			return ((base / p).string());
		}

		// Triangle area as in Triangle::_updateCache (Triangle.cpp:94-112).
		double	triangleArea(const Triangle& t)
		{
			const dvec3 n = cross(t.v1 - t.v0, t.v2 - t.v0);
			const double ls = dot(n, n);
			if (ls <= 1e-24)
			{
				return (0.0);
			}
			return (std::sqrt(ls) / 2.0);
		}

		dvec3	normalizedOrZero(const dvec3& n)
		{
			// OBJReader.cpp:57-70 (identity transform branch)
			const double ls = dot(n, n);
			if (ls <= 1e-24)
			{
				return {0.0, 0.0, 0.0};
			}
			if (std::fabs(ls - 1.0) <= 1e-12)
			{
				return (n);
			}
			const double len = std::sqrt(ls);
			return (n * (1.0 / len));
		}

		// --- materials: measured presets, glass dispersion and the material builder ---

		// MeasuredMaterials tables (MeasuredMaterials.cpp:53-150): conductor eta / k per RGB
		// channel, glass nd / Abbe vd with Sellmeier B / C (micrometres^2) where known. The
		// rendered colours are defined by these numbers, so they are the same numbers.
		struct	ConductorPreset { const char* name; dvec3 eta, k; };
		struct	GlassPreset { const char* name; double nd, vd; bool sellmeier; dvec3 b, c; };

		const ConductorPreset	kConductors[] = {
			{"aluminum", {1.44, 0.97, 0.62}, {7.38, 6.40, 5.30}},
			{"aluminium", {1.44, 0.97, 0.62}, {7.38, 6.40, 5.30}},
			{"copper", {0.27, 0.68, 1.34}, {3.57, 2.62, 2.29}},
			{"cu", {0.27, 0.68, 1.34}, {3.57, 2.62, 2.29}},
			{"gold", {0.17, 0.47, 1.54}, {3.14, 2.37, 1.91}},
			{"au", {0.17, 0.47, 1.54}, {3.14, 2.37, 1.91}},
			{"silver", {0.14, 0.16, 0.13}, {4.10, 3.14, 2.30}},
			{"ag", {0.14, 0.16, 0.13}, {4.10, 3.14, 2.30}},
			{"iron", {2.91, 2.95, 2.82}, {3.09, 2.93, 2.81}},
			{"fe", {2.91, 2.95, 2.82}, {3.09, 2.93, 2.81}},
			{"nickel", {1.92, 1.83, 1.62}, {3.67, 3.38, 3.04}},
			{"ni", {1.92, 1.83, 1.62}, {3.67, 3.38, 3.04}},
			{"chromium", {3.21, 3.18, 2.99}, {3.30, 3.32, 3.33}},
			{"chrome", {3.21, 3.18, 2.99}, {3.30, 3.32, 3.33}},
		};

		const GlassPreset	kGlasses[] = {
			{"bk7", 1.5168, 64.17, true, {1.03961212, 0.231792344, 1.01046945}, {0.00600069867, 0.0200179144, 103.560653}},
			{"borosilicate", 1.474, 65.0, false, {}, {}},
			{"fusedsilica", 1.4585, 67.82, true, {0.6961663, 0.4079426, 0.8974794}, {0.00467914826, 0.0135120631, 97.9340025}},
			{"silica", 1.4585, 67.82, true, {0.6961663, 0.4079426, 0.8974794}, {0.00467914826, 0.0135120631, 97.9340025}},
			{"water", 1.333, 55.8, false, {}, {}},
			{"diamond", 2.417, 55.3, false, {}, {}},
			{"sapphire", 1.773, 72.2, false, {}, {}},
		};

		// MeasuredMaterials volume table (MeasuredMaterials.cpp:139-152): scattering and
		// absorption coefficients (1/m) per RGB channel and the phase anisotropy.
		struct	VolumePreset { const char* name; dvec3 sigmaS, sigmaA; double g; };

		const VolumePreset	kVolumes[] = {
			{"clearair", {3.8e-6, 13.5e-6, 33.1e-6}, {0.0, 0.0, 0.0}, 0.0},
			{"air", {3.8e-6, 13.5e-6, 33.1e-6}, {0.0, 0.0, 0.0}, 0.0},
			{"haze", {7.5e-4, 8.5e-4, 1.0e-3}, {1.0e-4, 1.0e-4, 1.2e-4}, 0.55},
			{"mist", {0.025, 0.027, 0.030}, {0.001, 0.001, 0.001}, 0.72},
			{"fog", {0.050, 0.052, 0.055}, {0.002, 0.002, 0.002}, 0.15},
			{"smoke", {0.16, 0.15, 0.14}, {0.36, 0.34, 0.32}, 0.35},
			{"cloud", {2.0, 2.0, 2.0}, {0.02, 0.02, 0.02}, 0.85},
		};

		// preset names compare lower-case without '_', '-' or spaces (MeasuredMaterials.cpp:36-51)
		std::string	presetKey(const std::string& name)
		{
			std::string r;
			for (char ch : name)
			{
				const unsigned char u = static_cast<unsigned char>(ch);
				if (ch == '_' || ch == '-' || std::isspace(u)) continue;
				r.push_back(static_cast<char>(std::tolower(u)));
			}
			return (r);
		}

		template <typename T, std::size_t N>
		const T&	findPreset(const T (&table)[N], const std::string& name, const char* what)
		{
			const std::string key = presetKey(name);
			for (const T& t : table)
			{
				if (presetKey(t.name) == key) return (t);
			}
			std::string names;
			for (const T& t : table) names += (names.empty() ? "" : ", ") + std::string(t.name);
			throw std::invalid_argument(std::string("Unknown ") + what + " preset '" + name + "'. Available presets: " + names + ".");
		}

		// parseNanometers (Materials.cpp:95-115): "587.56" or "587.56nm"
		double	parseNanometers(std::string value, const std::string& label)
		{
			value = trim(value);
			const std::string l = lower(value);
			if (l.size() > 2 && l.compare(l.size() - 2, 2, "nm") == 0) value = trim(value.substr(0, value.size() - 2));
			std::size_t used = 0;
			const double nm = std::stod(value, &used);
			if (used != value.size() || !std::isfinite(nm) || nm <= 0.0)
			{
				throw std::runtime_error(label + " must be a finite positive wavelength in nm.");
			}
			return (nm);
		}

		// refractiveIndexFromSellmeier (MeasuredMaterials.cpp:254-283):
		// n^2 = 1 + sum_i B_i l^2 / (l^2 - C_i), l in micrometres
		double	sellmeierIndex(const dvec3& b, const dvec3& c, double nm)
		{
			if (!std::isfinite(nm) || nm <= 0.0) throw std::invalid_argument("Sellmeier wavelength must be finite and positive.");
			const double um = nm * 0.001;
			const double l2 = um * um;
			double n2 = 1.0;
			for (int i = 0; i < 3; i++)
			{
				if (!std::isfinite(b[i]) || !std::isfinite(c[i]) || c[i] < 0.0)
				{
					throw std::invalid_argument("Sellmeier coefficients must be finite, with non-negative C terms.");
				}
				const double den = l2 - c[i];
				if (std::fabs(den) <= 1e-12) throw std::invalid_argument("Sellmeier wavelength hits a resonance denominator.");
				n2 += (b[i] * l2) / den;
			}
			if (!std::isfinite(n2) || n2 <= 0.0) throw std::invalid_argument("Sellmeier coefficients produced an invalid refractive index.");
			return (std::sqrt(n2));
		}

		// refractiveIndexFromAbbe (MeasuredMaterials.cpp:285-304): the two-term Cauchy law
		// n(l) = A + B / l^2 through n_d at the d line with (n_d - 1) / V_d = n_F - n_C.
		double	abbeIndex(double nd, double vd, double nm)
		{
			auto positive = [](double v, const char* label) {
				if (!std::isfinite(v) || v <= 0.0) throw std::invalid_argument(std::string(label) + " must be finite and positive.");
			};
			positive(nd, "Abbe refractive index");
			positive(vd, "Abbe number");
			positive(nm, "Abbe wavelength");
			const double lF = 486.1327, lD = 587.5618, lC = 656.2725;
			const double dispersion = (nd - 1.0) / vd;
			const double b = dispersion / ((1.0 / (lF * lF)) - (1.0 / (lC * lC)));
			const double a = nd - (b / (lD * lD));
			return (a + (b / (nm * nm)));
		}

		// The keys of one material block, as parseMaterialProperty collects them
		// (MaterialBuilder, Materials.cpp:26-75).
		struct	MaterialKeys
		{
			std::string	type;
			dvec3	color = {0.6, 0.6, 0.6};
			dvec3	emissionColor = {1.0, 1.0, 1.0};
			bool	hasEmissionColor = false, hasProperties = false;
			double	roughness = 0.0, fuzz = -1.0;
			double	ior = 1.52;	// RI_GLASS (RefractiveIndexes.hpp:3)
			bool	hasIor = false;
			double	metallic = 0.0, transmission = 0.0, clearcoat = 0.0, clearcoatRoughness = 0.03, sheen = 0.0;
			double	glossyWeight = 0.5, alphaKey = 1.0;
			bool	hasSubsurface = false, hasSubsurfaceRadius = false, hasSubsurfaceScale = false;
			bool	hasSubsurfaceColor = false, hasSubsurfaceMethod = false;
			double	subsurface = 0.0, subsurfaceScale = 0.001;
			dvec3	subsurfaceRadius = {1.0, 1.0, 1.0}, subsurfaceColor = {1.0, 1.0, 1.0};
			bool	subsurfaceThin = false;
			std::string	subsurfaceProfile;
			bool	anySubsurface() const
			{
				return hasSubsurface || hasSubsurfaceRadius || hasSubsurfaceScale || hasSubsurfaceColor || hasSubsurfaceMethod
					|| !subsurfaceProfile.empty();
			}
			dvec3	glossyColor = {1.0, 1.0, 1.0};
			bool	hasGlossyColor = false;
			std::optional<double>	radiance, lum;
			double	anisotropy = 0.0;
			std::optional<dvec3>	eta, extinction, absorption, transmittance;
			std::string	materialPreset, conductorPreset, glassPreset;
			std::optional<double>	abbe;
			std::optional<dvec3>	sellmeierB, sellmeierC;
			double	iorWavelength = 587.5618;
			bool	hasIorWavelength = false;
			double	transmittanceDistance = 1.0;
			bool	hasTransmittanceDistance = false;
		};

		// materialRefractiveIndex (Materials.cpp:251-320)
		double	materialIor(const MaterialKeys& k, const std::string& type)
		{
			const GlassPreset* glass = nullptr;
			if (!k.glassPreset.empty()) glass = &findPreset(kGlasses, k.glassPreset, "glass");
			else if (!k.materialPreset.empty() && (type == "dielectric" || type == "principled")) glass = &findPreset(kGlasses, k.materialPreset, "glass");
			if (glass && k.hasIor) throw std::runtime_error("Material defines both glass preset and explicit ior.");
			if (glass && k.abbe) throw std::runtime_error("Material defines both glass preset and explicit Abbe number.");
			if (glass && (k.sellmeierB || k.sellmeierC)) throw std::runtime_error("Material defines both glass preset and explicit Sellmeier coefficients.");
			if (k.abbe && (k.sellmeierB || k.sellmeierC)) throw std::runtime_error("Material defines both Abbe and Sellmeier dispersion.");
			if (k.sellmeierB || k.sellmeierC)
			{
				if (!k.sellmeierB || !k.sellmeierC) throw std::runtime_error("Sellmeier dispersion requires both sellmeier_b and sellmeier_c.");
				return (sellmeierIndex(*k.sellmeierB, *k.sellmeierC, k.iorWavelength));
			}
			if (glass && glass->sellmeier) return (sellmeierIndex(glass->b, glass->c, k.iorWavelength));
			if (k.abbe) return (abbeIndex(glass ? glass->nd : k.ior, *k.abbe, k.iorWavelength));
			if (glass)
			{
				if (k.hasIorWavelength && glass->vd > 0.0) return (abbeIndex(glass->nd, glass->vd, k.iorWavelength));
				if (k.hasIorWavelength) throw std::runtime_error("ior_wavelength requires Sellmeier or Abbe dispersion.");
				return (glass->nd);
			}
			if (k.hasIorWavelength) throw std::runtime_error("ior_wavelength requires a glass preset, Abbe number, or Sellmeier coefficients.");
			return (k.ior);
		}

		// absorptionFromTransmittance (Dielectric.cpp:42-53, Principled.cpp): sigma_a = -ln(T) / d,
		// T = 0 -> T_MAX
		dvec3	absorptionOf(const MaterialKeys& k, const char* who)
		{
			if (k.absorption && k.transmittance) throw std::runtime_error(std::string(who) + " material defines both absorption and transmittance.");
			if (k.hasTransmittanceDistance && !k.transmittance) throw std::runtime_error(std::string(who) + " attenuation distance requires transmittance.");
			if (k.absorption)
			{
				const dvec3 a = *k.absorption;
				for (int i = 0; i < 3; i++)
				{
					if (!std::isfinite(a[i]) || a[i] < 0.0) throw std::invalid_argument(std::string(who) + " absorption coefficient must have finite non-negative channels.");
				}
				return (a);
			}
			if (k.transmittance)
			{
				if (!std::isfinite(k.transmittanceDistance) || k.transmittanceDistance <= 0.0)
				{
					throw std::invalid_argument(std::string(who) + " transmittance distance must be finite and positive.");
				}
				dvec3 a{};
				for (int i = 0; i < 3; i++)
				{
					const double t = (*k.transmittance)[i];
					if (!std::isfinite(t) || t < 0.0 || t > 1.0) throw std::invalid_argument(std::string(who) + " transmittance channels must be finite and in [0,1].");
					a[i] = t <= 0.0 ? std::numeric_limits<double>::max() : -std::log(t) / k.transmittanceDistance;
				}
				return (a);
			}
			return {0.0, 0.0, 0.0};
		}

		double	unitValue(double v, const std::string& label)
		{
			if (!std::isfinite(v) || v < 0.0 || v > 1.0) throw std::invalid_argument(label + " must be finite and in [0,1].");
			return (v);
		}

		void	conductorFrom(Material& m, const dvec3& eta, const dvec3& k, double roughness)
		{
			// Metal(eta, k, roughness), Metal.cpp:82-90, 118-125
			for (int i = 0; i < 3; i++)
			{
				if (!std::isfinite(eta[i]) || eta[i] <= 0.0) throw std::invalid_argument("Conductor eta must have finite positive channels.");
				if (!std::isfinite(k[i]) || k[i] < 0.0) throw std::invalid_argument("Conductor extinction coefficient must have finite non-negative channels.");
			}
			m.kind = MAT_METAL;
			m.color = {1.0, 1.0, 1.0};
			m.conductor = true;
			m.eta = eta;
			m.k = k;
			m.roughness = unitValue(roughness, "Metal roughness");
		}

		// buildMaterial (Materials.cpp:757-889) with buildMetalMaterial (:660-691),
		// buildDielectricMaterial (:323-346), buildPrincipledMaterial (:693-755) and
		// buildEmissiveMaterial (:210-224, LightUnits.cpp:36-73)
		Material	buildMaterial(const MaterialKeys& k)
		{
			if (!k.hasProperties) throw std::runtime_error("Material block ended before a material was defined.");
			const std::string type = k.type.empty() ? "lambertian" : k.type;
			if (!k.materialPreset.empty() && type != "metal" && type != "dielectric" && type != "principled")
			{
				throw std::runtime_error("Material preset is only valid for metal, dielectric, or principled materials.");
			}
			if (!k.conductorPreset.empty() && type != "metal") throw std::runtime_error("Conductor preset is only valid for metal materials.");
			if ((!k.glassPreset.empty() || k.abbe || k.sellmeierB || k.sellmeierC || k.hasIorWavelength) && type != "dielectric" && type != "principled")
			{
				throw std::runtime_error("Glass dispersion properties are only valid for dielectric or principled materials.");
			}
			if (k.anySubsurface() && type != "principled") throw std::runtime_error("Subsurface properties are only valid for principled materials.");
			Material m;
			m.color = k.color;
			if (type == "principled")
			{
				if (k.eta || k.extinction) throw std::runtime_error("Principled conductor metals use metallic color; use type=metal for eta/k.");
				m.absorption = absorptionOf(k, "Principled");
				m.kind = MAT_PRINCIPLED;
				m.metallic = unitValue(k.metallic, "Principled metallic");
				m.roughness = unitValue(k.roughness, "Principled roughness");
				m.transmission = unitValue(std::max(k.transmission, 1.0 - k.alphaKey), "Principled transmission");
				const double ior = materialIor(k, "principled");
				if (!std::isfinite(ior) || ior <= 0.0) throw std::invalid_argument("Principled refractive index must be finite and positive.");
				m.ior = ior;
				m.clearcoat = unitValue(k.clearcoat, "Principled clearcoat");
				m.clearcoatRoughness = unitValue(k.clearcoatRoughness, "Principled clearcoat roughness");
				m.sheen = unitValue(k.sheen, "Principled sheen");
				// the subsurface profile, then method, radius, scale, colour and weight
				// (Materials.cpp:727-757 over Principled.cpp:39-88)
				auto finiteNonNegative = [](const dvec3& c, const std::string& what) {
					for (double v : c)
					{
						if (!std::isfinite(v) || v < 0.0) throw std::invalid_argument(what + " must have finite non-negative channels.");
					}
				};
				if (!k.subsurfaceProfile.empty())
				{
					if (k.subsurfaceProfile != "skin") throw std::runtime_error("Unknown subsurface profile: " + k.subsurfaceProfile);
					m.subsurfaceThin = false;
					m.subsurfaceRadius = {1.0, 0.35, 0.18};
					m.subsurfaceScale = 0.0012;
					m.subsurfaceColor = {1.0, 0.42, 0.28};
					if (m.subsurface <= 0.0) m.subsurface = 0.5;
				}
				if (k.hasSubsurfaceMethod) m.subsurfaceThin = k.subsurfaceThin;
				if (k.hasSubsurfaceRadius)
				{
					finiteNonNegative(k.subsurfaceRadius, "Principled subsurface radius");
					if (!(k.subsurfaceRadius[0] > 0.0 || k.subsurfaceRadius[1] > 0.0 || k.subsurfaceRadius[2] > 0.0))
					{
						throw std::invalid_argument("Principled subsurface radius must have at least one positive channel.");
					}
					m.subsurfaceRadius = k.subsurfaceRadius;
				}
				if (k.hasSubsurfaceScale)
				{
					if (!std::isfinite(k.subsurfaceScale) || k.subsurfaceScale <= 0.0)
					{
						throw std::invalid_argument("Principled subsurface scale must be finite and positive.");
					}
					m.subsurfaceScale = k.subsurfaceScale;
				}
				if (k.hasSubsurfaceColor)
				{
					finiteNonNegative(k.subsurfaceColor, "Principled subsurface color");
					m.subsurfaceColor = k.subsurfaceColor;
				}
				if (k.hasSubsurface) m.subsurface = unitValue(k.subsurface, "Principled subsurface");
				return (m);
			}
			if (type == "glossy")
			{
				m.kind = MAT_GLOSSY;
				m.roughness = unitValue(k.roughness, "Glossy roughness");
				return (m);
			}
			if (type == "diffuse_glossy" || type == "diffuseglossy")
			{
				// DiffuseGlossy(color, glossyColor or color, weight, roughness)
				m.kind = MAT_DIFFUSE_GLOSSY;
				m.glossyColor = k.hasGlossyColor ? k.glossyColor : k.color;
				m.glossyWeight = unitValue(k.glossyWeight, "Diffuse-glossy weight");
				m.roughness = unitValue(k.roughness, "Diffuse-glossy roughness");
				return (m);
			}
			if (type == "lambertian")
			{
				m.kind = MAT_LAMBERTIAN;
				return (m);
			}
			if (type == "metal")
			{
				const double r = k.fuzz >= 0.0 ? k.fuzz : k.roughness;
				const std::string preset = !k.conductorPreset.empty() ? k.conductorPreset : k.materialPreset;
				if (!preset.empty())
				{
					if (k.eta || k.extinction) throw std::runtime_error("Metal material defines both conductor preset and explicit eta/k.");
					const ConductorPreset& c = findPreset(kConductors, preset, "conductor");
					conductorFrom(m, c.eta, c.k, r);
					return (m);
				}
				if (k.eta || k.extinction)
				{
					if (!k.eta || !k.extinction) throw std::runtime_error("Conductor metal requires both eta and k.");
					conductorFrom(m, *k.eta, *k.extinction, r);
					return (m);
				}
				m.kind = MAT_METAL;
				m.roughness = unitValue(r, "Metal roughness");
				return (m);
			}
			if (type == "dielectric")
			{
				m.absorption = absorptionOf(k, "Dielectric");
				const double ior = materialIor(k, "dielectric");
				if (!std::isfinite(ior) || ior <= 0.0) throw std::invalid_argument("Dielectric refractive index must be finite and positive.");
				m.kind = MAT_DIELECTRIC;
				m.ior = ior;
				m.roughness = unitValue(k.roughness, "Dielectric roughness");
				return (m);
			}
			if (type == "emissive")
			{
				const dvec3 c = k.hasEmissionColor ? k.emissionColor : k.color;
				double target;
				if (k.radiance) target = *k.radiance;
				else if (k.lum) target = *k.lum / 683.0;
				else throw std::runtime_error("Emissive material requires radiance or luminance.");
				if (!std::isfinite(target) || target < 0.0) throw std::invalid_argument("Light radiance must be finite and non-negative.");
				m.kind = MAT_EMISSIVE;
				if (target == 0.0)
				{
					m.color = {0.0, 0.0, 0.0};
				}
				else
				{
					const double cl = luminance(c);
					if (!(cl > 0.0)) throw std::runtime_error("Light color must have positive luminance for physical units.");
					const double f = target / cl;
					m.color = {c[0] * f, c[1] * f, c[2] * f};
				}
				return (m);
			}
			if (type == "isotropic")
			{
				m.kind = MAT_ISOTROPIC;
				return (m);
			}
			if (type == "volume" || type == "phase" || type == "henyey_greenstein" || type == "henyeygreenstein" || type == "hg")
			{
				// HenyeyGreenstein.cpp:10-18
				if (!std::isfinite(k.anisotropy) || k.anisotropy < -0.99 || k.anisotropy > 0.99)
				{
					throw std::invalid_argument("Volume anisotropy must be finite and between -0.99 and 0.99.");
				}
				m.kind = MAT_HG;
				m.anisotropy = k.anisotropy;
				return (m);
			}
			throw std::runtime_error("Unknown material type: " + type);
		}

		// parseDirectMaterialLine (Materials.cpp:117-196): `lambertian=(r,g,b)`,
		// `metal=(r,g,b),fuzz`, `glossy=(r,g,b),roughness`, `dielectric=(r,g,b)`,
		// `isotropic=(r,g,b)`, `hg=(r,g,b),g`; false when the line is none of them
		bool	directMaterialLine(const std::string& line, Material& m)
		{
			const std::string l = lower(trim(line));
			double r = 0, g = 0, b = 0, x = 0;
			auto starts = [&](const char* p) { return l.rfind(p, 0) == 0; };
			m = Material{};
			if (starts("lambertian="))
			{
				if (std::sscanf(l.c_str(), "lambertian=(%lf,%lf,%lf)", &r, &g, &b) != 3) throw std::runtime_error("Invalid lambertian material: " + line);
				m.kind = MAT_LAMBERTIAN;
			}
			else if (starts("metal="))
			{
				if (std::sscanf(l.c_str(), "metal=(%lf,%lf,%lf),%lf", &r, &g, &b, &x) != 4) throw std::runtime_error("Invalid metal material: " + line);
				m.kind = MAT_METAL;
				m.roughness = unitValue(x, "Metal roughness");
			}
			else if (starts("glossy="))
			{
				if (std::sscanf(l.c_str(), "glossy=(%lf,%lf,%lf),%lf", &r, &g, &b, &x) != 4) throw std::runtime_error("Invalid glossy material: " + line);
				m.kind = MAT_GLOSSY;
				m.roughness = unitValue(x, "Glossy roughness");
			}
			else if (starts("dielectric="))
			{
				if (std::sscanf(l.c_str(), "dielectric=(%lf,%lf,%lf)", &r, &g, &b) != 3) throw std::runtime_error("Invalid dielectric material: " + line);
				m.kind = MAT_DIELECTRIC;
				m.ior = 1.52;
			}
			else if (starts("emissive="))
			{
				throw std::runtime_error("Direct emissive material syntax is not physical. Use type=emissive with radiance or luminance.");
			}
			else if (starts("isotropic="))
			{
				if (std::sscanf(l.c_str(), "isotropic=(%lf,%lf,%lf)", &r, &g, &b) != 3) throw std::runtime_error("Invalid isotropic material: " + line);
				m.kind = MAT_ISOTROPIC;
			}
			else if (starts("henyey_greenstein=") || starts("henyeygreenstein=") || starts("hg="))
			{
				if (std::sscanf(l.c_str(), "%*[^=]=(%lf,%lf,%lf),%lf", &r, &g, &b, &x) != 4) throw std::runtime_error("Invalid Henyey-Greenstein material: " + line);
				if (!std::isfinite(x) || x < -0.99 || x > 0.99) throw std::invalid_argument("Volume anisotropy must be finite and between -0.99 and 0.99.");
				m.kind = MAT_HG;
				m.anisotropy = x;
			}
			else
			{
				return (false);
			}
			m.color = {r, g, b};
			return (true);
		}

		// A light block's physical quantity (PhysicalLightUnits, Objects.cpp:420-431): at most
		// one per block.
		struct	LightQuantity
		{
			enum Kind { NONE, RADIANCE, LUMINANCE, POWER, LUMENS, INTENSITY, CANDELA };
			Kind	kind = NONE;
			double	value = 0.0;
		};

		// parseSurfaceLightUnit / parseSphericalLightUnit (Objects.cpp:514-581): false when
		// `key` names none of the block's quantities (intensities only where `spherical`).
		bool	lightQuantity(LightQuantity& q, const std::string& key, const std::string& value, bool spherical)
		{
			LightQuantity::Kind k = LightQuantity::NONE;
			if (key == "radiance" || key == "surface_radiance" || key == "surfaceradiance") k = LightQuantity::RADIANCE;
			else if (key == "luminance" || key == "nits" || key == "cd_m2" || key == "cdm2") k = LightQuantity::LUMINANCE;
			else if (key == "power" || key == "watts" || key == "radiant_power" || key == "radiantpower") k = LightQuantity::POWER;
			else if (key == "lumens" || key == "lumen" || key == "lm" || key == "luminous_flux" || key == "luminousflux") k = LightQuantity::LUMENS;
			else if (spherical && (key == "radiant_intensity" || key == "radiantintensity" || key == "w_sr" || key == "wsr")) k = LightQuantity::INTENSITY;
			else if (spherical && (key == "candela" || key == "cd" || key == "luminous_intensity" || key == "luminousintensity")) k = LightQuantity::CANDELA;
			else return (false);
			if (q.kind != LightQuantity::NONE)
			{
				throw std::runtime_error("Light block defines multiple physical light quantities.");
			}
			q.kind = k;
			q.value = std::stod(value);
			return (true);
		}

		// surfaceLightEmission (Objects.cpp:603-630) over LightUnits.cpp:36-99: the colour
		// scaled so that its luminance is the radiance the quantity gives an emitter of
		// `area` square metres (luminous units divided by 683 lm/W).
		dvec3	lightEmission(const dvec3& color, const LightQuantity& q, double area, const char* missing)
		{
			auto nonNegative = [](double v, const std::string& what) {
				if (!std::isfinite(v) || v < 0.0) throw std::invalid_argument(what + " must be finite and non-negative.");
			};
			auto positiveArea = [](double a, const std::string& what) {
				if (!std::isfinite(a) || a <= 0.0) throw std::invalid_argument(what + " must be finite and positive.");
			};
			double radiance = 0.0;
			switch (q.kind)
			{
				case LightQuantity::RADIANCE:
					radiance = q.value;
					break;
				case LightQuantity::LUMINANCE:
					nonNegative(q.value, "Surface luminance");
					radiance = q.value / 683.0;
					break;
				case LightQuantity::POWER:
					nonNegative(q.value, "Radiant power");
					positiveArea(area, "Emitter area");
					radiance = q.value / (kPi * area);
					break;
				case LightQuantity::LUMENS:
				{
					nonNegative(q.value, "Luminous flux");
					positiveArea(area, "Emitter area");
					const double nits = q.value / (kPi * area);
					nonNegative(nits, "Surface luminance");
					radiance = nits / 683.0;
					break;
				}
				case LightQuantity::INTENSITY:
					nonNegative(q.value, "Radiant intensity");
					positiveArea(area, "Emitter surface area");
					radiance = 4.0 * q.value / area;
					break;
				case LightQuantity::CANDELA:
				{
					nonNegative(q.value, "Luminous intensity");
					positiveArea(area, "Emitter surface area");
					const double nits = 4.0 * q.value / area;
					nonNegative(nits, "Surface luminance");
					radiance = nits / 683.0;
					break;
				}
				default:
					throw std::runtime_error(missing);
			}
			nonNegative(radiance, "Surface radiance");
			// scaleToLuminance (LightUnits.cpp:36-52)
			nonNegative(color[0], "Light color red channel");
			nonNegative(color[1], "Light color green channel");
			nonNegative(color[2], "Light color blue channel");
			if (radiance == 0.0)
			{
				return dvec3{0.0, 0.0, 0.0};
			}
			const double lum = luminance(color);
			if (!std::isfinite(lum) || lum <= 0.0)
			{
				throw std::invalid_argument("Light color must have positive luminance for physical units.");
			}
			const double f = radiance / lum;
			return dvec3{color[0] * f, color[1] * f, color[2] * f};
		}

		// "visible=0|1" of sphere blocks and sphere lights (Objects.cpp:924-932, 1127-1135)
		bool	visibleFlag(const std::string& value, const std::string& what)
		{
			int v = 0;
			if (std::sscanf(value.c_str(), "%d", &v) != 1 || (v != 0 && v != 1))
			{
				throw std::runtime_error(what + " visible must be 0 or 1.");
			}
			return (v != 0);
		}

		struct	Reader
		{
			std::ifstream	in;
			std::filesystem::path	base;
			Scene	scene;
			std::map<std::string, std::uint32_t>	materialIndex;
			std::map<std::string, std::uint32_t>	meshByPath;
			std::map<std::string, std::string>	meshPaths;	// [meshes]: name -> resolved file
			ColorContext	colors;

			bool	next(std::string& line)
			{
				if (!std::getline(in, line))
				{
					return (false);
				}
				if (!line.empty() && line.back() == '\r')
				{
					line.pop_back();
				}
				return (true);
			}

			void	settings()
			{
				std::string line;
				Settings& s = scene.settings;
				while (next(line))
				{
					if (line.empty())
					{
						break;
					}
					if (line[0] == '#')
					{
						continue;
					}
					std::string key, value;
					if (!splitAssignment(line, key, value))
					{
						throw std::runtime_error("Invalid setting: " + line);
					}
					if (key == "resolution")
					{
						long long w, h;
						if (std::sscanf(value.c_str(), "%lld,%lld", &w, &h) != 2 || w <= 0 || h <= 0)
						{
							throw std::runtime_error("Invalid resolution setting. Use resolution=WIDTH,HEIGHT.");
						}
						s.width = static_cast<std::uint32_t>(w);
						s.height = static_cast<std::uint32_t>(h);
					}
					else if (key == "samples")
					{
						s.samples = static_cast<std::uint32_t>(std::max(0.0, parseDouble(value, "samples setting")));
					}
					else if (key == "maxlightbounces")
					{
						s.maxBounces = static_cast<std::int32_t>(parseDouble(value, "maxlightbounces setting"));
					}
					else if (key == "adaptive" || key == "adaptivesampling" || key == "adaptive_sampling")
					{
						// Settings.cpp:210-224: an integer, 0 or 1
						int v = 0;
						if (std::sscanf(value.c_str(), "%d", &v) != 1)
						{
							throw std::runtime_error("Invalid adaptive setting. Use adaptive=0 or adaptive=1.");
						}
						if (v != 0 && v != 1)
						{
							throw std::runtime_error("Invalid adaptive setting. Use 0 or 1.");
						}
						s.adaptive = v != 0;
					}
					else if (key == "adaptiveminsamples" || key == "adaptive_min_samples")
					{
						// Settings.cpp:225-237, Scene.cpp:193-200
						int v = 0;
						if (std::sscanf(value.c_str(), "%d", &v) != 1)
						{
							throw std::runtime_error("Invalid adaptive minimum samples setting. Use adaptiveminsamples=N.");
						}
						if (v <= 0)
						{
							throw std::runtime_error("Adaptive minimum samples must be positive.");
						}
						s.adaptiveMinSamples = static_cast<std::uint32_t>(v);
					}
					else if (key == "adaptivethreshold" || key == "adaptive_threshold")
					{
						// Settings.cpp:238-249, Scene.cpp:221-228
						double v = 0.0;
						if (std::sscanf(value.c_str(), "%lf", &v) != 1)
						{
							throw std::runtime_error("Invalid adaptive threshold setting. Use adaptivethreshold=F.");
						}
						if (!std::isfinite(v) || v <= 0.0)
						{
							throw std::runtime_error("Adaptive threshold must be positive.");
						}
						s.adaptiveThreshold = v;
					}
					else if (key == "adaptivecheckinterval" || key == "adaptive_check_interval")
					{
						// Settings.cpp:250-261, Scene.cpp:207-214
						int v = 0;
						if (std::sscanf(value.c_str(), "%d", &v) != 1)
						{
							throw std::runtime_error("Invalid adaptive check interval setting. Use adaptivecheckinterval=N.");
						}
						if (v <= 0)
						{
							throw std::runtime_error("Adaptive check interval must be positive.");
						}
						s.adaptiveCheckInterval = static_cast<std::uint32_t>(v);
					}
					else if (key == "denoise")
					{
						s.denoise = parseDouble(value, "denoise setting") != 0.0;
					}
					else if (key == "distanceblueness")
					{
						// Settings.cpp:646-655: 0 or 1; the renderer never reads it (ColorHelper.cpp:1155)
						int v = 0;
						if (std::sscanf(value.c_str(), "%d", &v) != 1 || (v != 0 && v != 1))
						{
							throw std::runtime_error("Invalid distanceblueness setting. Use distanceblueness=0 or distanceblueness=1.");
						}
					}
					else if (key == "outputfilename")
					{
						// the output file is the caller's (image writers are outside the device path)
					}
					else if (key == "view_transform")
					{
						// parseViewTransform (Settings.cpp:36-57)
						const std::string v = lower(trim(value));
						if (v == "standard") s.viewTransform = 0;
						else if (v == "agx") s.viewTransform = 1;
						else if (v == "aces") s.viewTransform = 2;
						else if (v == "raw") s.viewTransform = 3;
						else
						{
							throw std::runtime_error("Invalid view_transform setting. Use standard, agx, aces, or raw.");
						}
					}
					else if (key == "bloom")
					{
						// Settings.cpp:281-291: an integer, 0 or 1
						int v = 0;
						if (std::sscanf(value.c_str(), "%d", &v) != 1)
						{
							throw std::runtime_error("Invalid bloom setting. Use bloom=0 or bloom=1.");
						}
						if (v != 0 && v != 1)
						{
							throw std::runtime_error("Invalid bloom setting. Use 0 or 1.");
						}
						s.bloom = v != 0;
					}
					else if (key == "exposure")
					{
						// Settings.cpp:292-301, Scene.cpp:261-268
						double v = 0.0;
						if (std::sscanf(value.c_str(), "%lf", &v) != 1)
						{
							throw std::runtime_error("Invalid exposure setting. Use exposure=F.");
						}
						if (!std::isfinite(v) || !std::isfinite(std::pow(2.0, v)))
						{
							throw std::runtime_error("Exposure must be finite.");
						}
						s.exposure = v;
					}
					else if (key == "photographic_exposure" || key == "photographicexposure" || key == "camera_exposure"
						|| key == "cameraexposure")
					{
						// applyPhotographicExposure (Settings.cpp:102-117), Scene.cpp:270-286:
						// exposure = log2(shutter * (ISO / 100) / N^2)
						double n = 0.0, t = 0.0, iso = 0.0;
						char tail = 0;
						if (std::sscanf(value.c_str(), "%lf , %lf , %lf %c", &n, &t, &iso, &tail) != 3)
						{
							throw std::runtime_error("Invalid photographic exposure setting. Use photographic_exposure=F_NUMBER,SHUTTER_SECONDS,ISO.");
						}
						if (!std::isfinite(n) || n <= 0.0 || !std::isfinite(t) || t <= 0.0 || !std::isfinite(iso) || iso <= 0.0)
						{
							throw std::runtime_error("Photographic exposure values must be finite and positive.");
						}
						const double v = std::log2((t * (iso / 100.0)) / (n * n));
						if (!std::isfinite(v) || !std::isfinite(std::pow(2.0, v)))
						{
							throw std::runtime_error("Exposure must be finite.");
						}
						s.exposure = v;
					}
					else if (key == "contrast")
					{
						// Settings.cpp:314-323, Scene.cpp:293-300
						double v = 0.0;
						if (std::sscanf(value.c_str(), "%lf", &v) != 1)
						{
							throw std::runtime_error("Invalid contrast setting. Use contrast=F.");
						}
						if (!std::isfinite(v) || v < 0.0)
						{
							throw std::runtime_error("Contrast must be non-negative.");
						}
						s.contrast = v;
					}
					else if (key == "sky")
					{
						// Settings.cpp:440-470
						const std::string v = lower(value);
						if (v == "none") s.sky = SKY_NONE;
						else if (v == "environment") s.sky = SKY_ENVIRONMENT;
						else if (v == "linear") s.sky = SKY_LINEAR;
						else if (v == "atmosphere") s.sky = SKY_ATMOSPHERE;
						else throw std::runtime_error("Unknown sky setting: " + value);
						s.skySet = true;
					}
					else if (key == "atmosphere")
					{
						// Settings.cpp:612-628 with the checks of Atmosphere's constructor and
						// setters (Atmosphere.cpp:84-113, 186-300)
						if (s.sky != SKY_ATMOSPHERE)
						{
							throw std::runtime_error("Atmosphere settings require sky=atmosphere.");
						}
						double sun, earth, outer, hr, hm, stars;
						int samples, lightSamples;
						if (std::sscanf(lower(value).c_str(), "%lf,%lf,%lf,%lf,%lf,%d,%d,%lf", &sun, &earth, &outer, &hr, &hm, &samples,
								&lightSamples, &stars) != 8)
						{
							throw std::runtime_error("Invalid atmosphere setting. Use atmosphere=SUN,EARTH_RADIUS,ATMOSPHERE_RADIUS,HR,HM,SAMPLES,LIGHT_SAMPLES,STARS.");
						}
						if (!std::isfinite(sun)) throw std::runtime_error("Sun angle must be finite.");
						if (!std::isfinite(earth) || earth <= 0.0) throw std::runtime_error("Earth radius must be positive.");
						if (!std::isfinite(outer) || outer <= earth)
						{
							throw std::runtime_error("Atmosphere radius must be larger than Earth radius.");
						}
						if (!std::isfinite(hr) || hr <= 0.0) throw std::runtime_error("Atmosphere HR must be positive.");
						if (!std::isfinite(hm) || hm <= 0.0) throw std::runtime_error("Atmosphere HM must be positive.");
						if (samples <= 0) throw std::runtime_error("Atmosphere sample count must be positive.");
						if (lightSamples <= 0) throw std::runtime_error("Atmosphere light sample count must be positive.");
						if (!std::isfinite(stars) || stars < 0.0) throw std::runtime_error("Stars brightness must be non-negative.");
						// The device marches samples x lightSamples steps per miss ray inside one
						// kernel launch; the reference (CPU) only slows down, a GPU launch of that
						// length would outlast any watchdog.
						if (samples > kMaxAtmosphereSamples || lightSamples > kMaxAtmosphereSamples)
						{
							throw std::runtime_error("Unsupported for the device path: atmosphere SAMPLES and LIGHT_SAMPLES above "
								+ std::to_string(kMaxAtmosphereSamples));
						}
						s.sunAngle = sun;
						s.earthRadius = earth;
						s.atmosphereRadius = outer;
						s.atmosphereHR = hr;
						s.atmosphereHM = hm;
						s.atmosphereSamples = samples;
						s.atmosphereLightSamples = lightSamples;
						s.starsBrightness = stars;
					}
					else if (key == "atmosphere_sun_scale" || key == "atmospheresunscale" || key == "atmosphere_sun_multiplier"
						|| key == "atmospheresunmultiplier")
					{
						// Settings.cpp:629-643, Atmosphere.cpp:207-214
						const double v = parseDouble(value, "Atmosphere sun scale");
						if (!std::isfinite(v) || v < 0.0)
						{
							throw std::runtime_error("Atmosphere sun radiance scale must be finite and non-negative.");
						}
						s.sunRadianceScale = v;
					}
					else if (key == "background" || key == "backgroundcolor" || key == "background_color")
					{
						s.background = parseColor(value, "background color");
					}
					else if (key == "environment" || key == "environmentmap" || key == "environment_map")
					{
						std::string path = value;
						std::string rotation;
						const std::size_t comma = value.find(',');
						if (comma != std::string::npos)
						{
							path = trim(value.substr(0, comma));
							rotation = trim(value.substr(comma + 1));
						}
						const double strength = scene.environment.strength;
						const bool lighting = scene.environment.lighting;
						scene.environment = readHDREnvironment(resolvePath(base, path));
						scene.environment.strength = strength;
						scene.environment.lighting = lighting;
						if (!rotation.empty())
						{
							scene.environment.rotation = parseDouble(rotation, "environment rotation");
						}
						scene.hasEnvironment = true;
						if (!s.skySet)
						{
							s.sky = SKY_ENVIRONMENT;
						}
					}
					else if (key == "environment_scale" || key == "environmentscale")
					{
						scene.environment.strength = parseDouble(value, "environment scale");
					}
					else if (key == "environment_lighting" || key == "environmentlighting")
					{
						scene.environment.lighting = parseDouble(value, "environment_lighting") != 0.0;
					}
					else if (key == "environmentrotation" || key == "environment_rotation")
					{
						scene.environment.rotation = parseDouble(value, "environment rotation");
					}
					else if (key == "meters_per_unit" || key == "metersperunit")
					{
						s.metersPerUnit = parseDouble(value, "meters_per_unit");
					}
					else
					{
						throw std::runtime_error("Unsupported setting for the device path: " + line);
					}
				}
			}

			// One `material NAME { ... }` block: readBraceMaterialBlock / parseMaterialProperty /
			// buildMaterial (SceneFile/Materials.cpp:368-640, 757-889). Keys are collected, then
			// the material is built with the reference's checks; measured presets and glass
			// dispersion are evaluated here, so the device sees constants only.
			void	materialBlock(const std::string& name)
			{
				std::string line;
				MaterialKeys k;
				std::optional<Material> direct;
				auto emissionQuantity = [&](std::optional<double>& dst, const std::string& value) {
					// assignSingleEmissionQuantity (Materials.cpp:198-208)
					if (k.radiance || k.lum)
					{
						throw std::runtime_error("Material block defines multiple emission unit quantities.");
					}
					dst = std::stod(value);
				};
				while (next(line))
				{
					const std::string t = trim(line);
					if (t.empty() || t[0] == '#')
					{
						continue;
					}
					if (t == "}")
					{
						Material m = direct ? *direct : buildMaterial(k);
						m.name = name;
						if (materialIndex.count(name))
						{
							throw std::runtime_error("Duplicate material name: " + name);
						}
						materialIndex[name] = static_cast<std::uint32_t>(scene.materials.size());
						scene.materials.push_back(m);
						return;
					}
					Material dm;
					if (directMaterialLine(t, dm))
					{
						if (direct || k.hasProperties)
						{
							throw std::runtime_error("Material block defines more than one material.");
						}
						direct = dm;
						continue;
					}
					std::string key, value;
					if (!splitAssignment(t, key, value) || key.empty() || value.empty())
					{
						throw std::runtime_error("Invalid material property: " + line);
					}
					if (direct)
					{
						throw std::runtime_error("Material block mixes direct material syntax with property syntax.");
					}
					k.hasProperties = true;
					if (key == "type") k.type = lower(value);
					else if (key == "preset" || key == "material_preset" || key == "materialpreset") k.materialPreset = value;
					else if (key == "color" || key == "basecolor" || key == "base_color") k.color = parseColor(value, "material color");
					else if (key == "emission" || key == "emissioncolor" || key == "emission_color")
					{
						k.emissionColor = parseColor(value, "emission color");
						k.hasEmissionColor = true;
					}
					else if (key == "fuzz") k.fuzz = std::stod(value);
					else if (key == "roughness") k.roughness = std::stod(value);
					else if (key == "metallic") k.metallic = std::stod(value);
					else if (key == "eta" || key == "conductor_eta") k.eta = parseColor(value, key);
					else if (key == "metal_preset" || key == "metalpreset" || key == "conductor" || key == "conductor_preset"
						|| key == "conductorpreset") k.conductorPreset = value;
					else if (key == "k" || key == "extinction" || key == "extinctioncoefficient" || key == "extinction_coefficient"
						|| key == "conductor_k") k.extinction = parseColor(value, key);
					else if (key == "transmission") k.transmission = std::stod(value);
					else if (key == "clearcoat" || key == "clear_coat" || key == "coat") k.clearcoat = std::stod(value);
					else if (key == "clearcoatroughness" || key == "clearcoat_roughness" || key == "clear_coat_roughness"
						|| key == "coatroughness" || key == "coat_roughness") k.clearcoatRoughness = std::stod(value);
					else if (key == "sheen") k.sheen = std::stod(value);
					else if (key == "subsurface" || key == "subsurface_weight" || key == "subsurfaceweight" || key == "sss")
					{
						k.subsurface = std::stod(value);
						k.hasSubsurface = true;
					}
					else if (key == "subsurface_radius" || key == "subsurfaceradius" || key == "sss_radius" || key == "sssradius")
					{
						k.subsurfaceRadius = parseColor(value, key);
						k.hasSubsurfaceRadius = true;
					}
					else if (key == "subsurface_scale" || key == "subsurfacescale" || key == "sss_scale" || key == "sssscale")
					{
						k.subsurfaceScale = std::stod(value);
						k.hasSubsurfaceScale = true;
					}
					else if (key == "subsurface_color" || key == "subsurfacecolor" || key == "sss_color" || key == "ssscolor")
					{
						k.subsurfaceColor = parseColor(value, key);
						k.hasSubsurfaceColor = true;
					}
					else if (key == "subsurface_method" || key == "subsurfacemethod" || key == "sss_method" || key == "sssmethod")
					{
						// parseSubsurfaceMethod (Materials.cpp:349-367)
						const std::string v = lower(trim(value));
						if (v == "burley" || v == "diffusion" || v == "normalized_diffusion" || v == "normalizeddiffusion") k.subsurfaceThin = false;
						else if (v == "thin" || v == "thin_translucent" || v == "thintranslucent") k.subsurfaceThin = true;
						else throw std::runtime_error("Unknown subsurface method: " + value);
						k.hasSubsurfaceMethod = true;
					}
					else if (key == "subsurface_profile" || key == "subsurfaceprofile" || key == "sss_profile" || key == "sssprofile")
					{
						k.subsurfaceProfile = lower(trim(value));
					}
					else if (key == "glossy_color" || key == "glossycolor" || key == "specular_color" || key == "specularcolor")
					{
						k.glossyColor = parseColor(value, "glossy color");
						k.hasGlossyColor = true;
					}
					else if (key == "glossy_weight" || key == "glossyweight" || key == "glossy_mix" || key == "glossymix")
						k.glossyWeight = std::stod(value);
					else if (key == "ior" || key == "refractiveindex" || key == "refractive_index")
					{
						k.ior = std::stod(value);
						k.hasIor = true;
					}
					else if (key == "glass" || key == "glass_preset" || key == "glasspreset" || key == "dielectric_preset"
						|| key == "dielectricpreset") k.glassPreset = value;
					else if (key == "abbe" || key == "abbe_number" || key == "abbenumber" || key == "vd") k.abbe = std::stod(value);
					else if (key == "sellmeier_b" || key == "sellmeierb") k.sellmeierB = parseTriple(value, key);
					else if (key == "sellmeier_c" || key == "sellmeierc") k.sellmeierC = parseTriple(value, key);
					else if (key == "ior_wavelength" || key == "iorwavelength" || key == "refractive_index_wavelength"
						|| key == "refractiveindexwavelength")
					{
						k.iorWavelength = parseNanometers(value, "IOR wavelength");
						k.hasIorWavelength = true;
					}
					else if (key == "absorption" || key == "absorptioncoefficient" || key == "absorption_coefficient" || key == "sigma_a"
						|| key == "sigmaa") k.absorption = parseColor(value, key);
					else if (key == "transmittance" || key == "transmittancecolor" || key == "transmittance_color" || key == "attenuation"
						|| key == "attenuation_color" || key == "attenuationcolor") k.transmittance = parseColor(value, key);
					else if (key == "attenuationdistance" || key == "attenuation_distance" || key == "absorptiondistance"
						|| key == "absorption_distance")
					{
						k.transmittanceDistance = std::stod(value);
						k.hasTransmittanceDistance = true;
					}
					else if (key == "alpha") k.alphaKey = std::stod(value);
					else if (key == "radiance" || key == "surfaceradiance" || key == "surface_radiance" || key == "emissionradiance"
						|| key == "emission_radiance") emissionQuantity(k.radiance, value);
					else if (key == "luminance" || key == "nits" || key == "cd_m2" || key == "cdm2" || key == "emissionluminance"
						|| key == "emission_luminance") emissionQuantity(k.lum, value);
					else if (key == "anisotropy" || key == "g") k.anisotropy = std::stod(value);
					else if (key == "texture" || key == "basecolortexture" || key == "base_color_texture" || key == "albedo")
					{
						throw std::runtime_error("Unsupported for the device path: material textures (" + key + ")");
					}
					else throw std::runtime_error("Unknown material property: " + line);
				}
				throw std::runtime_error("Material block '" + name + "' is missing a closing }.");
			}

			void	materials()
			{
				std::string line;
				while (next(line))
				{
					const std::string t = trim(line);
					if (t.empty())
					{
						break;
					}
					if (t[0] == '#')
					{
						continue;
					}
					std::string name;
					if (!namedBlock(t, "material", name))
					{
						throw std::runtime_error("Invalid material block header: " + line);
					}
					materialBlock(name);
				}
			}

			std::uint32_t	material(const std::string& name)
			{
				const auto it = materialIndex.find(name);
				if (it == materialIndex.end())
				{
					throw std::runtime_error("Unknown material: " + name);
				}
				return (it->second);
			}

			// addCameraBlock (SceneFile/Scene.cpp:49-214): keys are collected, then applied at
			// the block's end in the reference's order (focal length, sensor, f-number or
			// aperture, pinhole, focus, photographic exposure). The scene renders from its first
			// camera (Scene::getActiveCamera, Scene.cpp:586-594).
			void	cameraBlock()
			{
				std::string line;
				Camera c;
				c.position = {0.0, 0.0, 15.0};	// Camera.cpp:21-33 defaults
				std::optional<double> focal, sensorW, sensorH, fNumber, aperture, focus, shutter, iso;
				std::optional<bool> pinhole;
				auto positive = [](const std::string& v, const std::string& label, double scale) {
					const double x = parseDouble(v, label);
					if (!std::isfinite(x) || x <= 0.0)
					{
						throw std::runtime_error(label + " must be finite and positive.");
					}
					return x * scale;
				};
				while (next(line))
				{
					const std::string t = trim(line);
					if (t.empty() || t[0] == '#')
					{
						continue;
					}
					if (t == "}")
					{
						if (focal) c.focalLength = *focal;
						if (sensorW) c.sensorWidth = *sensorW;
						if (sensorH) c.sensorHeight = *sensorH;
						if (fNumber && aperture)
						{
							throw std::runtime_error("Camera block defines both f_stop and aperture_diameter.");
						}
						if (fNumber)
						{
							c.fNumber = *fNumber;
							c.pinhole = false;
						}
						if (aperture)
						{
							c.fNumber = c.focalLength / *aperture;
							c.pinhole = false;
						}
						if (pinhole)
						{
							if (*pinhole && aperture)
							{
								throw std::runtime_error("Camera block defines pinhole=1 and aperture_diameter.");
							}
							c.pinhole = *pinhole;
						}
						if (focus) c.focusDistance = *focus;
						if (shutter || iso)
						{
							if (!shutter || !iso)
							{
								throw std::runtime_error("Camera block photographic exposure requires shutter and iso.");
							}
							// Scene::setPhotographicExposure (Scene.cpp:270-286) with the camera's f-number
							const double v = std::log2((*shutter * (*iso / 100.0)) / (c.fNumber * c.fNumber));
							if (!std::isfinite(v) || !std::isfinite(std::pow(2.0, v)))
							{
								throw std::runtime_error("Exposure must be finite.");
							}
							scene.settings.exposure = v;
						}
						if (!scene.hasCamera)
						{
							scene.camera = c;
							scene.hasCamera = true;
						}
						return;
					}
					std::string key, value;
					if (!splitAssignment(t, key, value))
					{
						throw std::runtime_error("Invalid camera property: " + line);
					}
					if (key == "position") c.position = parseTriple(value, "camera position");
					else if (key == "direction") c.direction = parseTriple(value, "camera direction");
					else if (key == "up" || key == "updirection" || key == "up_direction") c.up = parseTriple(value, "camera up");
					else if (key == "focal_length") focal = positive(value, "Camera focal_length", 1.0);
					else if (key == "focal_length_mm") focal = positive(value, "Camera focal_length_mm", 0.001);
					else if (key == "sensor_width") sensorW = positive(value, "Camera sensor_width", 1.0);
					else if (key == "sensor_width_mm") sensorW = positive(value, "Camera sensor_width_mm", 0.001);
					else if (key == "sensor_height") sensorH = positive(value, "Camera sensor_height", 1.0);
					else if (key == "sensor_height_mm") sensorH = positive(value, "Camera sensor_height_mm", 0.001);
					else if (key == "f_stop" || key == "fstop" || key == "f_number" || key == "fnumber")
						fNumber = positive(value, "Camera f_stop", 1.0);
					else if (key == "aperture_diameter") aperture = positive(value, "Camera aperture_diameter", 1.0);
					else if (key == "aperture_diameter_mm") aperture = positive(value, "Camera aperture_diameter_mm", 0.001);
					else if (key == "pinhole")
					{
						const int v = std::stoi(value);
						if (v != 0 && v != 1)
						{
							throw std::runtime_error("Camera pinhole must be 0 or 1.");
						}
						pinhole = v == 1;
					}
					else if (key == "focus_distance") focus = positive(value, "Camera focus_distance", 1.0);
					else if (key == "shutter" || key == "shutter_seconds" || key == "shutterseconds" || key == "shutter_speed"
						|| key == "shutterspeed") shutter = positive(value, "Camera shutter", 1.0);
					else if (key == "iso") iso = positive(value, "Camera iso", 1.0);
					else throw std::runtime_error("Unknown camera property: " + line);
				}
				throw std::runtime_error("Camera block is missing a closing }.");
			}

			std::uint32_t	meshFor(const std::string& path)
			{
				const auto it = meshByPath.find(path);
				if (it != meshByPath.end())
				{
					return (it->second);
				}
				const std::uint32_t id = static_cast<std::uint32_t>(scene.meshes.size());
				scene.meshes.push_back(readObjMesh(path));
				meshByPath[path] = id;
				return (id);
			}

			void	addInstance(const std::string& path, const dvec3& position, const dvec3& rotation, const dvec3& scale, std::uint32_t mat)
			{
				// Objects.cpp:357-406: an invertible scale makes a MeshInstance over the
				// per-file Mesh; a degenerate scale bakes a Mesh (not handled here).
				if (!(std::fabs(scale[0]) > 1e-12 && std::fabs(scale[1]) > 1e-12 && std::fabs(scale[2]) > 1e-12))
				{
					throw std::runtime_error("Unsupported for the device path: object with a non-invertible scale");
				}
				Instance inst;
				inst.mesh = meshFor(path);
				inst.position = position;
				inst.rotationDegrees = rotation;
				inst.scale = scale;
				scene.instances.push_back(inst);
				scene.hittables.push_back(Hittable{HIT_INSTANCE, static_cast<std::uint32_t>(scene.instances.size() - 1), mat});
			}

			void	objectBlock(const std::string& name)
			{
				std::string line;
				std::string file, meshName;
				dvec3 position{}, rotation{}, scale = {1.0, 1.0, 1.0};
				std::int64_t mat = -1;
				while (next(line))
				{
					const std::string t = trim(line);
					if (t.empty() || t[0] == '#')
					{
						continue;
					}
					if (t == "}")
					{
						if (!file.empty() && !meshName.empty())
						{
							throw std::runtime_error("Object block '" + name + "' defines both mesh and file.");
						}
						if (file.empty() && meshName.empty())
						{
							throw std::runtime_error("Object block '" + name + "' ended before mesh or file was defined.");
						}
						std::string path;
						if (!meshName.empty())
						{
							const auto it = meshPaths.find(meshName);
							if (it == meshPaths.end())
							{
								throw std::runtime_error("Unknown mesh name: " + meshName);
							}
							path = it->second;
						}
						else
						{
							path = resolvePath(base, file);
						}
						if (mat < 0)
						{
							Material m;	// ObjectBlock default: Lambertian(0.6)
							mat = static_cast<std::int64_t>(scene.materials.size());
							scene.materials.push_back(m);
						}
						addInstance(path, position, rotation, scale, static_cast<std::uint32_t>(mat));
						return;
					}
					std::string key, value;
					if (!splitAssignment(t, key, value))
					{
						throw std::runtime_error("Invalid object property: " + t);
					}
					if (key == "mesh") meshName = value;
					else if (key == "file" || key == "path") file = value;
					else if (key == "position") position = parseTriple(value, "object position");
					else if (key == "rotation") rotation = parseTriple(value, "object rotation");
					else if (key == "scale") scale = parseTriple(value, "object scale");
					else if (key == "material") mat = material(value);
					else throw std::runtime_error("Unsupported object property for the device path: " + t);
				}
				throw std::runtime_error("Object block '" + name + "' is missing a closing }.");
			}

			// "(a),(b),...,rest" -> fixed number of triples then the remainder
			std::vector<std::string>	splitTopLevel(const std::string& s)
			{
				std::vector<std::string> parts;
				int depth = 0;
				std::string cur;
				for (char ch : s)
				{
					if (ch == '(') depth++;
					if (ch == ')') depth--;
					if (ch == ',' && depth == 0)
					{
						parts.push_back(cur);
						cur.clear();
						continue;
					}
					cur.push_back(ch);
				}
				parts.push_back(cur);
				return (parts);
			}

			// readPrimitiveMaterial (Objects.cpp:154-186): `material=NAME`, or `material[` and
			// then one direct material line (lambertian=(r,g,b) ...) closed by `]` on its own
			// line (_readMaterialSubSection, Materials.cpp:927-960).
			std::uint32_t	primitiveMaterial(const std::string& part, const std::string& line)
			{
				const std::string p = trim(part);
				if (lower(p).rfind("material", 0) == 0 && trim(p.substr(8)) == "[")
				{
					std::string l;
					std::optional<Material> m;
					while (next(l))
					{
						const std::string t = trim(l);
						if (t.empty() || t[0] == '#')
						{
							continue;
						}
						if (t == "]")
						{
							if (!m)
							{
								throw std::runtime_error("Material block ended before a material was defined.");
							}
							scene.materials.push_back(*m);
							return (static_cast<std::uint32_t>(scene.materials.size() - 1));
						}
						Material parsed;
						if (!directMaterialLine(t, parsed))
						{
							throw std::runtime_error("Unknown material line: " + l);
						}
						if (m)
						{
							throw std::runtime_error("Material block defines more than one material.");
						}
						m = parsed;
					}
					throw std::runtime_error("Material block is missing a closing ].");
				}
				std::string key, value;
				if (!splitAssignment(p, key, value) || key != "material")
				{
					throw std::runtime_error("Invalid primitive material. Use material[ for an inline material block or material=NAME for a named material: " + line);
				}
				if (value.empty() || value.find_first_of(" \t{}[]") != std::string::npos)
				{
					throw std::runtime_error("Invalid primitive material name: " + line);
				}
				return (material(value));
			}

			std::uint32_t	addHittable(HittableKind kind, std::size_t index, std::uint32_t mat)
			{
				scene.hittables.push_back(Hittable{kind, static_cast<std::uint32_t>(index), mat});
				return (static_cast<std::uint32_t>(scene.hittables.size() - 1));
			}

			// volume NAME { ... } (Objects.cpp:1224-1330, 690-760): a ConstantVolume over a
			// sphere or box boundary with an isotropic or Henyey-Greenstein phase function.
			void	volumeBlock(const std::string& name)
			{
				std::string line;
				std::string shape = "box";
				Volume v;
				v.size = {1.0, 1.0, 1.0};
				double density = 0.1;
				dvec3 color = {0.72, 0.78, 0.86};
				double g = 0.0;
				std::int64_t mat = -1;
				bool hasDensity = false, hasColor = false, hasAnisotropy = false;
				std::string preset;
				double densityScale = 1.0;
				std::optional<dvec3> sigmaS, sigmaA;
				while (next(line))
				{
					const std::string t = trim(line);
					if (t.empty() || t[0] == '#')
					{
						continue;
					}
					if (t == "}")
					{
						// applyMeasuredVolumeData (Objects.cpp:746-800): a preset or explicit
						// sigma_s / sigma_a (scaled by density_scale) set the density to the
						// luminance of the extinction and the colour to sigma_s / density
						// (MeasuredMaterials.cpp:306-330).
						auto measured = [&](dvec3 s, dvec3 a) {
							if (!std::isfinite(densityScale) || densityScale <= 0.0)
							{
								throw std::runtime_error("Volume '" + name + "' density_scale must be finite and positive.");
							}
							for (int i = 0; i < 3; i++)
							{
								s[i] = s[i] * densityScale;
								a[i] = a[i] * densityScale;
							}
							for (int i = 0; i < 3; i++)
							{
								if (!std::isfinite(s[i]) || s[i] < 0.0 || !std::isfinite(a[i]) || a[i] < 0.0)
								{
									throw std::invalid_argument("Volume scattering and absorption coefficients must be finite and non-negative.");
								}
							}
							density = luminance(dvec3{s[0] + a[0], s[1] + a[1], s[2] + a[2]});
							if (!std::isfinite(density) || density <= 0.0)
							{
								throw std::invalid_argument("Measured volume coefficients produce zero extinction.");
							}
							for (int i = 0; i < 3; i++) color[i] = std::min(1.0, s[i] / density);
						};
						if (!preset.empty())
						{
							if (sigmaS || sigmaA)
							{
								throw std::runtime_error("Volume '" + name + "' defines both preset and explicit sigma coefficients.");
							}
							if (hasDensity || hasColor || hasAnisotropy)
							{
								throw std::runtime_error("Volume '" + name + "' preset cannot be combined with density, color, or anisotropy overrides. Use density_scale.");
							}
							const VolumePreset& vp = findPreset(kVolumes, preset, "volume");
							g = vp.g;
							measured(vp.sigmaS, vp.sigmaA);
						}
						else if (sigmaS || sigmaA)
						{
							if (hasDensity || hasColor)
							{
								throw std::runtime_error("Volume '" + name + "' sigma coefficients cannot be combined with density or color.");
							}
							measured(sigmaS.value_or(dvec3{0.0, 0.0, 0.0}), sigmaA.value_or(dvec3{0.0, 0.0, 0.0}));
						}
						auto positive = [&](double x, const std::string& what) {
							if (!std::isfinite(x) || x <= 0.0)
							{
								throw std::runtime_error("Volume '" + name + "' " + what + " must be finite and positive.");
							}
						};
						if (mat < 0)
						{
							Material m;
							m.color = color;
							if (std::fabs(g) <= 1e-12)
							{
								m.kind = MAT_ISOTROPIC;
							}
							else
							{
								if (!std::isfinite(g) || g < -0.99 || g > 0.99)
								{
									throw std::invalid_argument("Volume anisotropy must be finite and between -0.99 and 0.99.");
								}
								m.kind = MAT_HG;
								m.anisotropy = g;
							}
							mat = static_cast<std::int64_t>(scene.materials.size());
							scene.materials.push_back(m);
						}
						if (shape == "sphere")
						{
							positive(v.radius, "radius");
							v.sphere = true;
						}
						else if (shape == "box" || shape == "cube")
						{
							positive(v.size[0], "width");
							positive(v.size[1], "height");
							positive(v.size[2], "depth");
							v.sphere = false;
						}
						else
						{
							throw std::runtime_error("Unknown volume shape: " + shape);
						}
						positive(density, "density");
						v.density = density * scene.settings.metersPerUnit;
						positive(v.density, "scene-unit density");
						scene.volumes.push_back(v);
						addHittable(HIT_VOLUME, scene.volumes.size() - 1, static_cast<std::uint32_t>(mat));
						return;
					}
					std::string key, value;
					if (!splitAssignment(t, key, value))
					{
						throw std::runtime_error("Invalid volume property: " + t);
					}
					if (key == "shape" || key == "type") shape = lower(value);
					else if (key == "position" || key == "center") v.position = parseTriple(value, "volume position");
					else if (key == "size" || key == "dimensions") v.size = parseTriple(value, "volume size");
					else if (key == "width") v.size[0] = parseDouble(value, key);
					else if (key == "height") v.size[1] = parseDouble(value, key);
					else if (key == "depth") v.size[2] = parseDouble(value, key);
					else if (key == "radius") v.radius = parseDouble(value, key);
					else if (key == "density" || key == "extinction" || key == "sigma_t")
					{
						density = parseDouble(value, key);
						hasDensity = true;
					}
					else if (key == "color" || key == "albedo" || key == "scatteringcolor" || key == "scattering_color")
					{
						color = parseColor(value, "volume color");
						hasColor = true;
					}
					else if (key == "anisotropy" || key == "g")
					{
						g = parseDouble(value, key);
						hasAnisotropy = true;
					}
					else if (key == "preset" || key == "volume_preset" || key == "volumepreset" || key == "medium") preset = value;
					else if (key == "density_scale" || key == "densityscale" || key == "coefficient_scale" || key == "coefficientscale")
						densityScale = parseDouble(value, key);
					else if (key == "sigma_s" || key == "sigmas" || key == "scattering" || key == "scattering_coefficient" || key == "scatteringcoefficient")
						sigmaS = parseColor(value, key);
					else if (key == "sigma_a" || key == "sigmaa" || key == "absorption" || key == "absorption_coefficient" || key == "absorptioncoefficient")
						sigmaA = parseColor(value, key);
					else if (key == "material")
					{
						mat = material(value);
						const MaterialKind k = scene.materials[static_cast<std::size_t>(mat)].kind;
						if (k != MAT_ISOTROPIC && k != MAT_HG)
						{
							throw std::runtime_error("Unsupported for the device path: volume with a surface material");
						}
					}
					else throw std::runtime_error("Unsupported volume property for the device path: " + t);
				}
				throw std::runtime_error("Volume '" + name + "' is missing a closing }.");
			}

			// directional_light NAME { ... } (Objects.cpp:1163-1222): a delta light with the
			// emission color scaled to the given irradiance / illuminance (LightUnits.cpp:101-111).
			void	directionalBlock(const std::string& name)
			{
				std::string line;
				DirectionalLight d;
				dvec3 direction = {0.0, -1.0, 0.0};
				dvec3 color = {1.0, 1.0, 1.0};
				bool hasColor = false;
				double irradiance = -1.0;
				double solar = -1.0;
				while (next(line))
				{
					const std::string t = trim(line);
					if (t.empty() || t[0] == '#')
					{
						continue;
					}
					if (t == "}")
					{
						if (dot(direction, direction) <= 0.0)
						{
							throw std::runtime_error("Directional light '" + name + "' direction must be non-zero.");
						}
						if (solar >= 0.0)
						{
							// solarDirectionalIrradiance (LightUnits.cpp:113-117): the solar
							// constant times the scale, in ColorScience::solar()'s colour unless
							// the light names one (Objects.cpp:1174-1189)
							if (irradiance >= 0.0) throw std::runtime_error("Light block defines multiple physical light quantities.");
							if (!hasColor) color = blackbodyColor(5778.0);
							irradiance = 1361.0 * solar;
							d.hasSunRadiance = true;
						}
						if (irradiance < 0.0)
						{
							throw std::runtime_error("Directional light must define irradiance, illuminance, or solar.");
						}
						Material m;
						m.kind = MAT_EMISSIVE;
						if (irradiance == 0.0)
						{
							m.color = {0.0, 0.0, 0.0};
						}
						else
						{
							const double cl = luminance(color);
							if (!(cl > 0.0))
							{
								throw std::invalid_argument("Light color must have positive luminance for physical units.");
							}
							const double f = irradiance / cl;
							m.color = {color[0] * f, color[1] * f, color[2] * f};
						}
						if (d.hasSunRadiance) d.sunRadiance = m.color;
						scene.materials.push_back(m);
						const double len = std::sqrt(dot(direction, direction));
						d.direction = direction * (1.0 / len);
						scene.directionals.push_back(d);
						addHittable(HIT_DIRECTIONAL, scene.directionals.size() - 1, static_cast<std::uint32_t>(scene.materials.size() - 1));
						return;
					}
					std::string key, value;
					if (!splitAssignment(t, key, value))
					{
						throw std::runtime_error("Invalid directional light property: " + t);
					}
					if (key == "direction") direction = parseTriple(value, "directional light direction");
					else if (key == "color")
					{
						color = parseColor(value, "directional light color");
						hasColor = true;
					}
					else if (key == "irradiance" || key == "w_m2" || key == "wm2" || key == "illuminance" || key == "lux" || key == "lx")
					{
						const double v = parseDouble(value, key);
						if (!(v >= 0.0))
						{
							throw std::invalid_argument("Directional irradiance must be finite and non-negative.");
						}
						if (irradiance >= 0.0 || solar >= 0.0) throw std::runtime_error("Light block defines multiple physical light quantities.");
						irradiance = (key == "irradiance" || key == "w_m2" || key == "wm2") ? v : v / 683.0;
					}
					else if (key == "solar" || key == "sun" || key == "solar_scale" || key == "solarscale")
					{
						const double v = parseDouble(value, key);
						if (!(v >= 0.0)) throw std::invalid_argument("Solar scale must be finite and non-negative.");
						if (irradiance >= 0.0 || solar >= 0.0) throw std::runtime_error("Light block defines multiple physical light quantities.");
						solar = v;
					}
					else throw std::runtime_error("Unsupported directional light property for the device path: " + t);
				}
				throw std::runtime_error("Directional light '" + name + "' is missing a closing }.");
			}

			// sphere NAME { ... } (Objects.cpp:880-951)
			void	sphereBlock(const std::string& name)
			{
				std::string line;
				Sphere sp;
				sp.radius = 1.0;
				std::int64_t mat = -1;
				bool visible = true;
				while (next(line))
				{
					const std::string t = trim(line);
					if (t.empty() || t[0] == '#')
					{
						continue;
					}
					if (t == "}")
					{
						if (!std::isfinite(sp.radius) || sp.radius <= 0.0)
						{
							throw std::runtime_error("Sphere object '" + name + "' radius must be finite and positive.");
						}
						sp.visible = visible;
						if (mat < 0)
						{
							scene.materials.push_back(Material{});	// SphereBlock default: Lambertian(0.6)
							mat = static_cast<std::int64_t>(scene.materials.size() - 1);
						}
						scene.spheres.push_back(sp);
						addHittable(HIT_SPHERE, scene.spheres.size() - 1, static_cast<std::uint32_t>(mat));
						return;
					}
					std::string key, value;
					if (!splitAssignment(t, key, value) || key.empty() || value.empty())
					{
						throw std::runtime_error("Invalid sphere property: " + t);
					}
					if (key == "position" || key == "center") sp.position = parseTriple(value, key);
					else if (key == "radius") sp.radius = std::stod(value);
					else if (key == "material") mat = material(value);
					else if (key == "visible") visible = visibleFlag(value, "Sphere object '" + name + "'");
					else if (key == "uv_projection" || key == "uvprojection" || key == "texture_projection" || key == "projection")
					{
						// parseSphereUVProjection (Objects.cpp:662-686): it places textures,
						// which the device path does not take, so it is only checked
						const std::string v = lower(value);
						if (v != "latlong" && v != "lat_long" && v != "longitude_latitude" && v != "equirectangular"
							&& v != "cube_cross" && v != "cubecross" && v != "cross" && v != "cube")
						{
							throw std::runtime_error("Unknown sphere uv projection: " + value);
						}
					}
					else throw std::runtime_error("Unknown sphere property: " + t);
				}
				throw std::runtime_error("Sphere object '" + name + "' is missing a closing }.");
			}

			// area_light NAME { ... } (Objects.cpp:955-1048): an emissive Rectangle facing
			// `normal`, its radiance from the block's physical quantity over w*h scene units^2.
			void	areaLightBlock(const std::string& name)
			{
				std::string line;
				Rectangle r;
				r.orientation = {0.0, -1.0, 0.0};
				dvec3 color = {1.0, 1.0, 1.0};
				LightQuantity q;
				while (next(line))
				{
					const std::string t = trim(line);
					if (t.empty() || t[0] == '#')
					{
						continue;
					}
					if (t == "}")
					{
						if (r.width <= 0.0 || r.height <= 0.0)
						{
							throw std::runtime_error("Area light '" + name + "' size must be positive.");
						}
						if (dot(r.orientation, r.orientation) <= 0.0)
						{
							throw std::runtime_error("Area light '" + name + "' normal must be non-zero.");
						}
						const double mpu = scene.settings.metersPerUnit;
						Material m;
						m.kind = MAT_EMISSIVE;
						m.color = lightEmission(color, q, r.width * r.height * mpu * mpu,
							"Surface light must define radiance, luminance, power, or lumens.");
						scene.materials.push_back(m);
						scene.rectangles.push_back(r);
						addHittable(HIT_RECTANGLE, scene.rectangles.size() - 1, static_cast<std::uint32_t>(scene.materials.size() - 1));
						return;
					}
					std::string key, value;
					if (!splitAssignment(t, key, value) || key.empty() || value.empty())
					{
						throw std::runtime_error("Invalid area light property: " + t);
					}
					if (key == "position") r.position = parseTriple(value, key);
					else if (key == "normal" || key == "direction") r.orientation = parseTriple(value, key);
					else if (key == "size")
					{
						// parseVector2Value (Objects.cpp:200-212)
						if (std::sscanf(trim(value).c_str(), "(%lf,%lf)", &r.width, &r.height) != 2)
						{
							throw std::runtime_error("Invalid size value. Use size=(x,y).");
						}
					}
					else if (key == "width") r.width = std::stod(value);
					else if (key == "height") r.height = std::stod(value);
					else if (key == "color") color = parseColor(value, key);
					else if (!lightQuantity(q, key, value, false)) throw std::runtime_error("Unknown area light property: " + t);
				}
				throw std::runtime_error("Area light '" + name + "' is missing a closing }.");
			}

			// sphere_light / point_light NAME { ... } (Objects.cpp:1050-1160): an emissive
			// Sphere, its radiance from the block's quantity over its 4 pi r^2 surface.
			void	sphereLightBlock(const std::string& name)
			{
				std::string line;
				Sphere sp;
				sp.radius = 0.1;
				dvec3 color = {1.0, 1.0, 1.0};
				LightQuantity q;
				bool visible = true;
				std::string iesPath;
				dvec3 iesDirection = {0.0, -1.0, 0.0};
				double iesRotation = 0.0;
				while (next(line))
				{
					const std::string t = trim(line);
					if (t.empty() || t[0] == '#')
					{
						continue;
					}
					if (t == "}")
					{
						if (sp.radius <= 0.0)
						{
							throw std::runtime_error("Sphere light '" + name + "' radius must be positive.");
						}
						sp.visible = visible;
						const double mpu = scene.settings.metersPerUnit;
						if (!iesPath.empty())
						{
							// the profile is read (and checked) before the emission, as the
							// reference builds it (Objects.cpp:1066-1089)
							const double rotation = iesRotation;
							if (!std::isfinite(rotation)) throw std::invalid_argument("IES profile rotation must be finite.");
							const double l2 = dot(iesDirection, iesDirection);
							scene.iesProfiles.push_back(readIesProfile(resolvePath(base, iesPath)));
							if (!std::isfinite(l2) || l2 <= 0.0) throw std::invalid_argument("IES profile direction must be non-zero.");
							sp.ies = static_cast<std::int32_t>(scene.iesProfiles.size() - 1);
							sp.iesDirection = iesDirection;
							sp.iesRotation = rotation;
							// sphericalLightEmission (Objects.cpp:632-643): without a unit of its
							// own the light emits the profile's luminous flux
							if (q.kind == LightQuantity::NONE)
							{
								q.kind = LightQuantity::LUMENS;
								q.value = scene.iesProfiles.back().totalLumens;
							}
						}
						Material m;
						m.kind = MAT_EMISSIVE;
						m.color = lightEmission(color, q, 4.0 * kPi * sp.radius * sp.radius * mpu * mpu,
							"Sphere light must define radiance, luminance, power, lumens, candela, radiant_intensity, or ies.");
						scene.materials.push_back(m);
						scene.spheres.push_back(sp);
						addHittable(HIT_SPHERE, scene.spheres.size() - 1, static_cast<std::uint32_t>(scene.materials.size() - 1));
						return;
					}
					std::string key, value;
					if (!splitAssignment(t, key, value) || key.empty() || value.empty())
					{
						throw std::runtime_error("Invalid sphere light property: " + t);
					}
					if (key == "position") sp.position = parseTriple(value, key);
					else if (key == "radius") sp.radius = std::stod(value);
					else if (key == "color") color = parseColor(value, key);
					else if (lightQuantity(q, key, value, true)) continue;
					else if (key == "ies" || key == "ies_profile" || key == "iesprofile" || key == "profile") iesPath = value;
					else if (key == "ies_direction" || key == "iesdirection" || key == "profile_direction" || key == "profiledirection")
						iesDirection = parseTriple(value, key);
					else if (key == "ies_rotation" || key == "iesrotation" || key == "profile_rotation" || key == "profilerotation")
						iesRotation = std::stod(value);
					else if (key == "visible") visible = visibleFlag(value, "Sphere light '" + name + "'");
					else throw std::runtime_error("Unknown sphere light property: " + t);
				}
				throw std::runtime_error("Sphere light '" + name + "' is missing a closing }.");
			}

			// addSceneObjectOrLightBlock (Objects.cpp:1338-1391): the named blocks that may
			// stand in [scene] or inside objects{ }.
			bool	namedSceneBlock(const std::string& t)
			{
				std::string name;
				if (namedBlock(t, "object", name)) objectBlock(name);
				else if (namedBlock(t, "sphere", name)) sphereBlock(name);
				else if (namedBlock(t, "area_light", name)) areaLightBlock(name);
				else if (namedBlock(t, "directional_light", name)) directionalBlock(name);
				else if (namedBlock(t, "volume", name)) volumeBlock(name);
				else if (namedBlock(t, "sphere_light", name) || namedBlock(t, "point_light", name)) sphereLightBlock(name);
				else
				{
					const std::string lt = lower(t);
					for (const char* k : {"object ", "sphere ", "area_light ", "directional_light ", "volume ", "sphere_light ", "point_light "})
					{
						if (lt.rfind(k, 0) == 0) throw std::runtime_error("Invalid scene block header: " + t);
					}
					return (false);
				}
				return (true);
			}

			void	objects()
			{
				std::string line;
				while (next(line))
				{
					const std::string t = trim(line);
					if (t.empty() || t[0] == '#')
					{
						continue;
					}
					if (t == "}")
					{
						return;
					}
					const std::string lt = lower(t);
					if (namedSceneBlock(t))
					{
						continue;
					}
					if (lt.rfind("triangle=", 0) == 0)
					{
						const auto p = splitTopLevel(t.substr(9));
						if (p.size() != 4)
						{
							throw std::runtime_error("Invalid triangle object: " + line);
						}
						Triangle tri;
						tri.v0 = parseTriple(p[0], "triangle vertex");
						tri.v1 = parseTriple(p[1], "triangle vertex");
						tri.v2 = parseTriple(p[2], "triangle vertex");
						scene.triangles.push_back(tri);
						scene.hittables.push_back(Hittable{HIT_TRIANGLE, static_cast<std::uint32_t>(scene.triangles.size() - 1), primitiveMaterial(p[3], line)});
						continue;
					}
					if (lt.rfind("rectangle=", 0) == 0)
					{
						const auto p = splitTopLevel(t.substr(10));
						if (p.size() != 5)
						{
							throw std::runtime_error("Unsupported rectangle form for the device path: " + line);
						}
						Rectangle r;
						r.position = parseTriple(p[0], "rectangle position");
						r.orientation = parseTriple(p[1], "rectangle orientation");
						r.width = parseDouble(p[2], "rectangle width");
						r.height = parseDouble(p[3], "rectangle height");
						scene.rectangles.push_back(r);
						scene.hittables.push_back(Hittable{HIT_RECTANGLE, static_cast<std::uint32_t>(scene.rectangles.size() - 1), primitiveMaterial(p[4], line)});
						continue;
					}
					if (lt.rfind("sphere=", 0) == 0)
					{
						// sphere=(x,y,z),r,material=NAME (Objects.cpp:1425-1448)
						const auto p = splitTopLevel(t.substr(7));
						if (p.size() != 3)
						{
							throw std::runtime_error("Invalid sphere object: " + line);
						}
						Sphere sp;
						sp.position = parseTriple(p[0], "sphere position");
						sp.radius = parseDouble(p[1], "sphere radius");
						scene.spheres.push_back(sp);
						addHittable(HIT_SPHERE, scene.spheres.size() - 1, primitiveMaterial(p[2], line));
						continue;
					}
					if (lt.rfind("cube=", 0) == 0)
					{
						// cube=(p),(o),w,h,d,material=NAME (Objects.cpp:1449-1474)
						const auto p = splitTopLevel(t.substr(5));
						if (p.size() != 6)
						{
							throw std::runtime_error("Invalid cube object: " + line);
						}
						Cube c;
						c.position = parseTriple(p[0], "cube position");
						(void)parseTriple(p[1], "cube orientation");
						c.width = parseDouble(p[2], "cube width");
						c.height = parseDouble(p[3], "cube height");
						c.depth = parseDouble(p[4], "cube depth");
						scene.cubes.push_back(c);
						addHittable(HIT_CUBE, scene.cubes.size() - 1, primitiveMaterial(p[5], line));
						continue;
					}
					if (lt.rfind("plane=", 0) == 0)
					{
						// plane=y,(o),material=NAME (Objects.cpp:1475-1496)
						const auto p = splitTopLevel(t.substr(6));
						if (p.size() != 3)
						{
							throw std::runtime_error("Invalid plane object: " + line);
						}
						Plane pl;
						pl.y = parseDouble(p[0], "plane y");
						pl.orientation = parseTriple(p[1], "plane orientation");
						scene.planes.push_back(pl);
						addHittable(HIT_PLANE, scene.planes.size() - 1, primitiveMaterial(p[2], line));
						continue;
					}
					if (lt.rfind("obj=", 0) == 0)
					{
						const auto p = splitTopLevel(t.substr(4));
						if (p.size() == 1)
						{
							Material m;
							scene.materials.push_back(m);
							addInstance(resolvePath(base, trim(p[0])), dvec3{}, dvec3{}, dvec3{1.0, 1.0, 1.0},
								static_cast<std::uint32_t>(scene.materials.size() - 1));
							continue;
						}
						if (p.size() == 3)
						{
							addInstance(resolvePath(base, trim(p[0])), parseTriple(p[1], "obj position"), dvec3{},
								dvec3{1.0, 1.0, 1.0}, primitiveMaterial(p[2], line));
							continue;
						}
						throw std::runtime_error("Invalid OBJ object: " + line);
					}
					throw std::runtime_error("Unsupported object line for the device path: " + t);
				}
				throw std::runtime_error("Objects section is missing a closing }.");
			}

			void	sceneSection()
			{
				std::string line;
				while (next(line))
				{
					const std::string t = trim(line);
					if (t.empty())
					{
						break;
					}
					if (t[0] == '#')
					{
						continue;
					}
					std::string name;
					if (namedBlock(t, "camera", name))
					{
						cameraBlock();
					}
					else if (lower(t) == "objects{" || lower(t) == "objects {")
					{
						objects();
					}
					else if (namedSceneBlock(t))
					{
					}
					else
					{
						throw std::runtime_error("Unsupported scene entry for the device path: " + t);
					}
				}
			}

			// [spectra] (Parsing.cpp:292-368): `reflectance NAME {` blocks of "nm value" lines,
			// up to the section's first empty line; each block becomes the colour
			// reflectance(NAME) stands for.
			void	spectra()
			{
				std::string line;
				while (next(line))
				{
					const std::string t = trim(line);
					if (t.empty())
					{
						return;
					}
					if (t[0] == '#')
					{
						continue;
					}
					std::string name;
					if (!namedBlock(t, "reflectance", name))
					{
						throw std::runtime_error("Invalid spectrum block header: " + line);
					}
					if (colors.spectra.count(name) != 0)
					{
						throw std::runtime_error("Duplicate spectrum name: " + name);
					}
					std::vector<std::pair<double, double>> samples;
					bool closed = false;
					while (!closed && next(line))
					{
						std::string clean = line.substr(0, line.find('#'));
						std::replace(clean.begin(), clean.end(), ',', ' ');
						std::replace(clean.begin(), clean.end(), ';', ' ');
						if (trim(clean) == "}")
						{
							try
							{
								colors.spectra[name] = reflectanceColor(samples);
							}
							catch (const std::exception& e)
							{
								throw std::runtime_error("Invalid spectrum '" + name + "': " + e.what());
							}
							closed = true;
							continue;
						}
						std::pair<double, double> sample;
						if (spectrumSample(line, sample, "'" + name + "'")) samples.push_back(sample);
					}
					if (!closed)
					{
						throw std::runtime_error("Spectrum block '" + name + "' is missing a closing }.");
					}
				}
			}

			// [meshes] (Meshes.cpp:4-74): `mesh NAME { file=PATH }` blocks naming OBJ files
			// that object blocks refer to with mesh=NAME.
			void	meshes()
			{
				std::string line;
				while (next(line))
				{
					const std::string t = trim(line);
					if (t.empty())
					{
						return;
					}
					if (t[0] == '#')
					{
						continue;
					}
					std::string name;
					if (!namedBlock(t, "mesh", name))
					{
						throw std::runtime_error("Invalid mesh block header: " + line);
					}
					if (meshPaths.count(name) != 0)
					{
						throw std::runtime_error("Duplicate mesh name: " + name);
					}
					std::string file;
					bool closed = false;
					while (!closed && next(line))
					{
						const std::string b = trim(line);
						if (b.empty() || b[0] == '#')
						{
							continue;
						}
						if (b == "}")
						{
							if (file.empty())
							{
								throw std::runtime_error("Mesh block '" + name + "' ended before file was defined.");
							}
							meshPaths[name] = resolvePath(base, file);
							closed = true;
							continue;
						}
						std::string key, value;
						if (!splitAssignment(b, key, value) || key.empty() || value.empty())
						{
							throw std::runtime_error("Invalid mesh property: " + b);
						}
						if (key == "file" || key == "path") file = value;
						else throw std::runtime_error("Unknown mesh property: " + b);
					}
					if (!closed)
					{
						throw std::runtime_error("Mesh block '" + name + "' is missing a closing }.");
					}
				}
			}

			void	run()
			{
				colors.base = base;
				struct Scope
				{
					const ColorContext* saved;
					explicit Scope(const ColorContext* c) : saved(tColorContext) { tColorContext = c; }
					~Scope() { tColorContext = saved; }
				} scope(&colors);
				std::string line;
				while (next(line))
				{
					const std::string t = trim(line);
					if (t.empty() || t[0] == '#')
					{
						continue;
					}
					const std::string lt = lower(t);
					if (lt.rfind("[settings]", 0) == 0) settings();
					else if (lt.rfind("[spectra]", 0) == 0) spectra();
					else if (lt.rfind("[materials]", 0) == 0) materials();
					else if (lt.rfind("[meshes]", 0) == 0) meshes();
					else if (lt.rfind("[scene]", 0) == 0) sceneSection();
					else throw std::runtime_error("Unsupported section for the device path: " + t);
				}
			}
		};
	}

	Scene	readSceneFile(const std::string& path)
	{
		Reader r;
		r.in.open(path);
		if (!r.in)
		{
			throw std::runtime_error("Scene file could not be opened: " + path);
		}
		r.base = std::filesystem::path(path).parent_path();
		r.run();
		if (!r.scene.hasCamera)
		{
			throw std::runtime_error("No camera on the Scene.");
		}
		return (std::move(r.scene));
	}

	Mesh	readObjMesh(const std::string& path)
	{
		FILE* f = std::fopen(path.c_str(), "rb");
		if (!f)
		{
			throw std::runtime_error("OBJ file could not be opened: " + path);
		}
		std::vector<dvec3> vertices, normals;
		struct FV { int v, n; };
		Mesh mesh;
		mesh.source = path;
		char buf[4096];
		std::vector<FV> face;
		auto resolve = [&](long idx, std::size_t count) -> int {
			long r = idx > 0 ? idx - 1 : static_cast<long>(count) + idx;
			if (idx == 0 || r < 0 || static_cast<std::size_t>(r) >= count)
			{
				throw std::runtime_error("OBJ index out of range in " + path);
			}
			return (static_cast<int>(r));
		};
		while (std::fgets(buf, sizeof(buf), f))
		{
			if (buf[0] == 'v' && buf[1] == ' ')
			{
				double x, y, z;
				if (std::sscanf(buf + 2, "%lf %lf %lf", &x, &y, &z) != 3)
				{
					std::fclose(f);
					throw std::runtime_error("Invalid OBJ vector line in " + path);
				}
				vertices.push_back({x, y, z});
			}
			else if (buf[0] == 'v' && buf[1] == 'n' && buf[2] == ' ')
			{
				double x, y, z;
				if (std::sscanf(buf + 3, "%lf %lf %lf", &x, &y, &z) != 3)
				{
					std::fclose(f);
					throw std::runtime_error("Invalid OBJ vector line in " + path);
				}
				normals.push_back({x, y, z});
			}
			else if (buf[0] == 'f' && buf[1] == ' ')
			{
				face.clear();
				const char* c = buf + 2;
				while (true)
				{
					while (std::isspace(static_cast<unsigned char>(*c))) c++;
					if (*c == '\0') break;
					char* end;
					const long vi = std::strtol(c, &end, 10);
					if (end == c)
					{
						std::fclose(f);
						throw std::runtime_error("Invalid OBJ face vertex index in " + path);
					}
					c = end;
					FV fv{resolve(vi, vertices.size()), -1};
					if (*c == '/')
					{
						c++;
						if (*c != '/' && !std::isspace(static_cast<unsigned char>(*c)) && *c)
						{
							std::strtol(c, &end, 10);	// texture index: unused on the device path
							c = end;
						}
						if (*c == '/')
						{
							c++;
							if (*c && !std::isspace(static_cast<unsigned char>(*c)))
							{
								fv.n = resolve(std::strtol(c, &end, 10), normals.size());
								c = end;
							}
						}
					}
					face.push_back(fv);
				}
				if (face.size() < 3)
				{
					std::fclose(f);
					throw std::runtime_error("OBJ face has fewer than three vertices in " + path);
				}
				for (std::size_t i = 1; i + 1 < face.size(); i++)
				{
					Triangle t;
					t.v0 = vertices[face[0].v];
					t.v1 = vertices[face[i].v];
					t.v2 = vertices[face[i + 1].v];
					if (face[0].n >= 0 && face[i].n >= 0 && face[i + 1].n >= 0)
					{
						t.n0 = normalizedOrZero(normals[face[0].n]);
						t.n1 = normalizedOrZero(normals[face[i].n]);
						t.n2 = normalizedOrZero(normals[face[i + 1].n]);
						t.hasNormals = dot(t.n0, t.n0) > 1e-24 && dot(t.n1, t.n1) > 1e-24 && dot(t.n2, t.n2) > 1e-24;
					}
					// OBJReader.cpp:436 skips area <= 1e-24; Mesh.cpp:369-379 drops area <= 0.
					if (triangleArea(t) > 1e-24)
					{
						mesh.triangles.push_back(t);
					}
				}
			}
		}
		std::fclose(f);
		return (mesh);
	}

	// IESProfile::load (IESProfile.cpp:133-232): an LM-63 photometric file. Lines up to
	// TILT=NONE (other tilt data is refused) are the header; then whitespace-separated
	// numbers: lamp count, lumens per lamp, candela multiplier, vertical and horizontal angle
	// counts, photometric type and seven more lamp fields, the vertical angles, the horizontal
	// angles, and the candela values horizontal-major.
	IesProfile	readIesProfile(const std::string& path)
	{
		std::ifstream in(path);
		if (!in) throw std::runtime_error("IES profile could not be opened: " + path);
		std::string line;
		bool tilt = false;
		while (std::getline(in, line))
		{
			const std::string t = lower(trim(line));
			if (t.rfind("tilt=", 0) == 0)
			{
				if (t != "tilt=none") throw std::runtime_error("IES profile uses unsupported tilt data: " + path);
				tilt = true;
				break;
			}
		}
		if (!tilt) throw std::runtime_error("IES profile is missing TILT=NONE: " + path);
		std::vector<double> v;
		while (std::getline(in, line))
		{
			std::istringstream ls(line);
			double x;
			while (ls >> x) v.push_back(x);
		}
		if (v.size() < 13) throw std::runtime_error("IES profile has incomplete lamp metadata: " + path);
		auto count = [&](double x, const char* what) {
			if (!std::isfinite(x) || x <= 0.0 || std::floor(x) != x) throw std::runtime_error(std::string(what) + " must be a positive integer.");
			return static_cast<std::size_t>(x);
		};
		const double lamps = v[0], lumensPerLamp = v[1], multiplier = v[2];
		const std::size_t nv = count(v[3], "IES vertical angle count");
		const std::size_t nh = count(v[4], "IES horizontal angle count");
		if (v.size() < 13 + nv + nh + nv * nh) throw std::runtime_error("IES profile has incomplete angle or candela data: " + path);
		if (!std::isfinite(lamps) || lamps <= 0.0) throw std::runtime_error("IES profile lamp count must be finite and positive: " + path);
		if (!std::isfinite(lumensPerLamp)) throw std::runtime_error("IES profile lumens per lamp must be finite: " + path);
		if (!std::isfinite(multiplier) || multiplier <= 0.0)
			throw std::runtime_error("IES profile candela multiplier must be finite and positive: " + path);
		IesProfile pr;
		std::size_t at = 13;
		pr.vertical.assign(v.begin() + static_cast<std::ptrdiff_t>(at), v.begin() + static_cast<std::ptrdiff_t>(at + nv));
		at += nv;
		pr.horizontal.assign(v.begin() + static_cast<std::ptrdiff_t>(at), v.begin() + static_cast<std::ptrdiff_t>(at + nh));
		at += nh;
		for (std::size_t i = 0; i < nv * nh; i++)
		{
			const double c = v[at + i] * multiplier;
			if (!std::isfinite(c) || c < 0.0) throw std::runtime_error("IES profile candela values must be finite and non-negative: " + path);
			pr.candela.push_back(c);
		}
		auto increasing = [](const std::vector<double>& a, const char* what) {
			for (std::size_t i = 0; i < a.size(); i++)
			{
				if (!std::isfinite(a[i])) throw std::runtime_error(std::string(what) + " contain a non-finite value.");
				if (i > 0 && a[i] <= a[i - 1]) throw std::runtime_error(std::string(what) + " must be strictly increasing.");
			}
		};
		increasing(pr.vertical, "IES vertical angles");
		increasing(pr.horizontal, "IES horizontal angles");
		// integrateTotalLumens (IESProfile.cpp:114-132): midpoint rule, 180 x 360 cells
		const double dTheta = kPi / 180.0, dPhi = (2.0 * kPi) / 360.0;
		double total = 0.0;
		for (int a = 0; a < 180; a++)
		{
			const double theta = (static_cast<double>(a) + 0.5) * dTheta;
			const double vdeg = theta * (180.0 / kPi);
			const double st = std::sin(theta);
			for (int b = 0; b < 360; b++)
			{
				const double hdeg = (static_cast<double>(b) + 0.5) * 360.0 / 360.0;
				total += iesCandelaAt(pr, vdeg, hdeg) * st * dTheta * dPhi;
			}
		}
		if (!std::isfinite(total) || total <= 0.0) throw std::runtime_error("IES profile has zero luminous flux: " + path);
		pr.totalLumens = total;
		pr.meanCandela = total / (4.0 * kPi);
		return (pr);
	}

	double	iesCandelaAt(const IesProfile& pr, double vdeg, double hdeg)
	{
		const std::vector<double>& V = pr.vertical;
		const std::vector<double>& H = pr.horizontal;
		if (V.empty() || H.empty() || pr.candela.empty() || !std::isfinite(vdeg)) return (0.0);
		if (vdeg < V.front() || vdeg > V.back()) return (0.0);
		std::size_t v0 = 0;
		while (v0 + 1 < V.size() && V[v0 + 1] < vdeg) v0++;
		const std::size_t v1 = std::min(v0 + 1, V.size() - 1);
		const double vSpan = V[v1] - V[v0];
		const double vt = vSpan > 0.0 ? (vdeg - V[v0]) / vSpan : 0.0;
		std::size_t h0 = 0, h1 = 0;
		double ht = 0.0;
		if (H.size() > 1)
		{
			// the angle in [0, 360), lifted by a turn below the first angle, held at a last
			// angle of 360 or more
			double h = 0.0;
			if (std::isfinite(hdeg))
			{
				h = std::fmod(hdeg, 360.0);
				if (h < 0.0) h += 360.0;
			}
			if (h < H.front()) h += 360.0;
			if (h > H.back() && H.back() >= 360.0) h = std::min(h, H.back());
			while (h0 + 1 < H.size() && H[h0 + 1] < h) h0++;
			if (h0 + 1 < H.size())
			{
				h1 = h0 + 1;
				const double span = H[h1] - H[h0];
				ht = span > 0.0 ? (h - H[h0]) / span : 0.0;
			}
			else
			{
				// past the last angle: towards the first one, a turn later
				h1 = 0;
				const double span = (H.front() + 360.0) - H[h0];
				ht = span > 0.0 ? (h - H[h0]) / span : 0.0;
			}
		}
		const std::size_t nv = V.size();
		auto lerp = [](double a, double b, double t) { return (a * (1.0 - t) + b * t); };
		const double c00 = pr.candela[h0 * nv + v0], c01 = pr.candela[h0 * nv + v1];
		const double c10 = pr.candela[h1 * nv + v0], c11 = pr.candela[h1 * nv + v1];
		return (lerp(lerp(c00, c01, vt), lerp(c10, c11, vt), ht));
	}

	Environment	readHDREnvironment(const std::string& path)
	{
		std::ifstream in(path, std::ios::binary);
		if (!in)
		{
			throw std::runtime_error("HDR environment map could not be opened: " + path);
		}
		std::string line;
		std::getline(in, line);
		if (line.rfind("#?RADIANCE", 0) != 0 && line.rfind("#?RGBE", 0) != 0)
		{
			throw std::runtime_error("Unsupported HDR environment header: " + path);
		}
		bool rgbe = false;
		while (std::getline(in, line))
		{
			if (line.empty() || line == "\r")
			{
				break;
			}
			if (line.rfind("FORMAT=", 0) == 0 && line.find("32-bit_rle_rgbe") != std::string::npos)
			{
				rgbe = true;
			}
		}
		if (!rgbe)
		{
			throw std::runtime_error("HDR environment map must use FORMAT=32-bit_rle_rgbe: " + path);
		}
		std::getline(in, line);
		char ya[8], xa[8];
		int h = 0, w = 0;
		if (std::sscanf(line.c_str(), "%7s %d %7s %d", ya, &h, xa, &w) != 4 || std::string(ya) != "-Y" || std::string(xa) != "+X" || w <= 0 || h <= 0)
		{
			throw std::runtime_error("Unsupported HDR environment resolution line (only -Y H +X W): " + line);
		}
		static const double M[3][3] = {	// ColorManagement.cpp:121-125, linear sRGB -> ACEScg
			{0.6131324224, 0.3395380158, 0.0474166960},
			{0.0701243808, 0.9163940113, 0.0134515239},
			{0.0205876575, 0.1095745716, 0.8697854040}
		};
		auto clampNonNegative = [](double v) { return (std::isfinite(v) && v > 0.0) ? v : 0.0; };
		auto decode = [&](const unsigned char* p) -> dvec3 {
			if (p[3] == 0)
			{
				return {0.0, 0.0, 0.0};
			}
			const double s = std::ldexp(1.0, static_cast<int>(p[3]) - (128 + 8));
			const double r = clampNonNegative(p[0] * s), g = clampNonNegative(p[1] * s), b = clampNonNegative(p[2] * s);
			dvec3 o;
			for (int i = 0; i < 3; i++)
			{
				o[i] = clampNonNegative(M[i][0] * r + M[i][1] * g + M[i][2] * b);
			}
			return (o);
		};
		Environment env;
		env.width = static_cast<std::uint32_t>(w);
		env.height = static_cast<std::uint32_t>(h);
		env.pixels.resize(static_cast<std::size_t>(w) * h);
		std::vector<unsigned char> scan(static_cast<std::size_t>(w) * 4);
		for (int y = 0; y < h; y++)
		{
			unsigned char head[4];
			if (!in.read(reinterpret_cast<char*>(head), 4))
			{
				throw std::runtime_error("Truncated HDR environment scanline header.");
			}
			const int encoded = (head[2] << 8) | head[3];
			if (w >= 8 && w <= 32767 && head[0] == 2 && head[1] == 2 && (head[2] & 0x80) == 0 && encoded == w)
			{
				for (int ch = 0; ch < 4; ch++)
				{
					int pos = 0;
					while (pos < w)
					{
						unsigned char code;
						if (!in.read(reinterpret_cast<char*>(&code), 1))
						{
							throw std::runtime_error("Truncated HDR environment RLE packet.");
						}
						if (code > 128)
						{
							const int count = code - 128;
							unsigned char value;
							if (pos + count > w || !in.read(reinterpret_cast<char*>(&value), 1))
							{
								throw std::runtime_error("Invalid HDR environment RLE run.");
							}
							for (int i = 0; i < count; i++) scan[(pos + i) * 4 + ch] = value;
							pos += count;
						}
						else
						{
							const int count = code;
							if (count == 0 || pos + count > w)
							{
								throw std::runtime_error("Invalid HDR environment RLE literal.");
							}
							for (int i = 0; i < count; i++)
							{
								unsigned char value;
								if (!in.read(reinterpret_cast<char*>(&value), 1))
								{
									throw std::runtime_error("Truncated HDR environment RLE literal.");
								}
								scan[(pos + i) * 4 + ch] = value;
							}
							pos += count;
						}
					}
				}
			}
			else
			{
				std::memcpy(scan.data(), head, 4);
				if (!in.read(reinterpret_cast<char*>(scan.data() + 4), static_cast<std::streamsize>(w - 1) * 4))
				{
					throw std::runtime_error("Truncated HDR environment scanline.");
				}
			}
			for (int x = 0; x < w; x++)
			{
				env.pixels[static_cast<std::size_t>(y) * w + x] = decode(&scan[x * 4]);
			}
		}
		return (env);
	}
}
