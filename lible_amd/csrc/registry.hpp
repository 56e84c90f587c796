// Table of quartet-kernel launchers, filled at load time by the instantiation units.
#pragma once

#include <hip/hip_runtime.h>

namespace lbx
{
    struct KArgs;
    using Launcher = void (*)(const KArgs &, hipStream_t);

    // Kernel variants of one class: one lane per quartet with the bra pairs on the lanes,
    // the same with the ket pairs on the lanes, one wave per quartet on MFMA, or one workgroup
    // per bra pair and block of ket pairs (ket transform on the VALU, bra transform on MFMA);
    // and those forms with R and both transforms in FP32 (J/K of the small-bound quartets of the
    // mixed-precision build; Boys, the epilogue and the accumulation stay FP64).
    enum Variant
    {
        VAR_VALU_BRA = 0,
        VAR_VALU_KET = 1,
        VAR_MFMA = 2,
        VAR_VALU_BRA32 = 3,
        VAR_VALU_KET32 = 4,
        VAR_MFMA32 = 5,
        VAR_BLOCK = 6,
        VAR_BLOCK32 = 7,
        N_VARIANTS = 8
    };

    // R and the transforms in FP32 (the small-bound tier of mixed-precision builds)
    constexpr bool isF32Variant(int v) { return (v >= VAR_VALU_BRA32 && v <= VAR_MFMA32) || v == VAR_BLOCK32; }

    // how a kernel walks its task table
    enum KernelKind
    {
        KIND_LANE = 0,    // 64 quartets per task, one per lane
        KIND_WAVE = 1,    // one quartet per task, one wave
        KIND_BLOCK = 2    // one bra pair x up to `width` ket pairs per task, one workgroup
    };

    struct KernelSlot
    {
        Launcher fn = nullptr;
        int width = 0;      // quartets per task: 64 (one lane each), 1 (one wave), or the ket block
        bool swap = false;  // lanes run over the ket class, the bra pair is wave-uniform
        int nc = 0;         // column-pass width of the VALU variants
        int kind = KIND_LANE;
    };

    // mode 0 J/K, 1 dense, 2 diagonal, 3 per-molecule sums; la, lc <= 5, lb, ld <= 3
    KernelSlot &kernelSlot(int mode, int la, int lb, int lc, int ld, int variant);
}
