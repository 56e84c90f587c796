// Table of quartet-kernel launchers, filled at load time by the instantiation units.
#pragma once

#include <hip/hip_runtime.h>

namespace lbx
{
    struct KArgs;
    using Launcher = void (*)(const KArgs &, hipStream_t);

    // Kernel variants of one class: one lane per quartet with the bra pairs on the lanes,
    // the same with the ket pairs on the lanes, or one wave per quartet on MFMA; and those
    // three forms with R and both transforms in FP32 (J/K of the small-bound quartets of the
    // mixed-precision build; Boys, the epilogue and the accumulation stay FP64).
    enum Variant
    {
        VAR_VALU_BRA = 0,
        VAR_VALU_KET = 1,
        VAR_MFMA = 2,
        N_FP64_VARIANTS = 3,
        VAR_VALU_BRA32 = 3,
        VAR_VALU_KET32 = 4,
        VAR_MFMA32 = 5,
        N_VARIANTS = 6
    };

    struct KernelSlot
    {
        Launcher fn = nullptr;
        int width = 0;      // quartets per task: 64 (one lane each) or 1 (one wave)
        bool swap = false;  // lanes run over the ket class, the bra pair is wave-uniform
        int nc = 0;         // column-pass width of the VALU variants
    };

    // mode 0 J/K, 1 dense, 2 diagonal, 3 per-molecule sums; la, lc <= 5, lb, ld <= 3
    KernelSlot &kernelSlot(int mode, int la, int lb, int lc, int ld, int variant);
}
