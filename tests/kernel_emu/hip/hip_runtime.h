#include "../hip_cpu_shim.hpp"
