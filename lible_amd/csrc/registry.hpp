// Table of quartet-kernel launchers, filled at load time by the instantiation units.
#pragma once

#include <hip/hip_runtime.h>

namespace lbx
{
    struct KArgs;
    using Launcher = void (*)(const KArgs &, hipStream_t);

    struct KernelSlot
    {
        Launcher fn = nullptr;
        int width = 0;   // quartets per task: 64 (one lane each) or 1 (one wave)
    };

    // mode 0 J/K, 1 dense, 2 diagonal; la, lc <= 5, lb, ld <= 3
    KernelSlot &kernelSlot(int mode, int la, int lb, int lc, int ld);
}
