// Library metadata entry points.
#include "common.h"

extern "C" const char* bd_version(void) { return "bubbledynamics-amd 0.1 (gfx950)"; }

extern "C" int bd_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
