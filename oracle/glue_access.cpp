// Test infrastructure: compiles the reference-side binding (luz_amd/glue/luzgpu_glue.cpp)
// against the reference's unmodified headers. The binding reads members the reference keeps
// private; a maintainer grants that with `friend struct luzgpu_glue::Access;` in those
// classes. Here the two defines below stand in for those friend lines (the standard headers
// are included first, so only the reference's declarations are affected).
#include <algorithm>
#include <atomic>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#define private public
#define protected public
#include "../luz_amd/glue/luzgpu_glue.cpp"
#undef private
#undef protected
