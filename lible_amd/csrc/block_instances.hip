// Instantiation unit of the block J/K kernels (quartet_block.hpp).  Compiled once per
// LBX_BUNIT: 0..9 the FP64 forms with bra = main pair class LBX_BUNIT (order (0,0),(1,0),(1,1),
// (2,0),...) against every ket class <= it, 10..19 the FP32 forms of the same class pairs.
#include "quartet_block.hpp"

#ifndef LBX_BUNIT
#error "LBX_BUNIT must be defined"
#endif

namespace lbx
{
    namespace
    {
        constexpr bool kF32Unit = LBX_BUNIT >= 10;
        constexpr int kPairUnit = kF32Unit ? LBX_BUNIT - 10 : LBX_BUNIT;

        constexpr int unitLa()
        {
            int la = 0;
            while (pairClassIdx(la + 1, 0) <= kPairUnit)
                la++;
            return la;
        }
        constexpr int PA = unitLa(), PB = kPairUnit - pairClassIdx(PA, 0);
        constexpr int CI = pairClassIdx(PA, PB);

        template <int LC, int LD>
        void reg()
        {
            if constexpr (pairClassIdx(LC, LD) <= CI)
                registerBlock<PA, PB, LC, LD, kF32Unit ? 2 : 1>();
        }

        struct Registrar
        {
            Registrar()
            {
                reg<0, 0>();
                reg<1, 0>();
                reg<1, 1>();
                reg<2, 0>();
                reg<2, 1>();
                reg<2, 2>();
                reg<3, 0>();
                reg<3, 1>();
                reg<3, 2>();
                reg<3, 3>();
            }
        } registrar;
    }
}
