// Instantiation unit of the quartet kernels.  Compiled once per LBX_UNIT:
//   0..9   bra = main pair class LBX_UNIT (order (0,0),(1,0),(1,1),(2,0),...): eri4 J/K (FP64
//          forms) and dense for every ket class <= bra class, the (ab|ab) diagonal, and eri3
//          (ab|c) against every auxiliary shell l = 0..5 not already covered;
//   10, 11 bra = single auxiliary shell l = 4, 5: eri2 (a|b) for b <= l and the (a|a) diagonal;
//   12..23 the FP32 forms of units 0..11: J/K and the batched RI sums.
// Splitting the instantiations lets the build compile the units in parallel.
#include "quartet.hpp"

#ifndef LBX_UNIT
#error "LBX_UNIT must be defined"
#endif

namespace lbx
{
    namespace
    {
        constexpr bool kF32Unit = LBX_UNIT >= 12;
        constexpr int kPairUnit = kF32Unit ? LBX_UNIT - 12 : LBX_UNIT;

        constexpr int unitLa()
        {
            constexpr int u = kPairUnit;
            if (u >= 10)
                return u - 6;   // 10 -> 4, 11 -> 5
            int la = 0;
            while (pairClassIdx(la + 1, 0) <= u)
                la++;
            return la;
        }
        constexpr int unitLb() { return kPairUnit >= 10 ? 0 : kPairUnit - pairClassIdx(unitLa(), 0); }

        constexpr int PA = unitLa(), PB = unitLb();
        constexpr int CI = pairClassIdx(PA, PB);

        template <int LC, int LD>
        void regEri4()
        {
            if constexpr (kPairUnit < 10 && pairClassIdx(LC, LD) <= CI)
            {
                if constexpr (kF32Unit)
                    registerQuartet<PA, PB, LC, LD, MODE_JK, 2>();
                else
                {
                    registerQuartet<PA, PB, LC, LD, MODE_JK, 1>();
                    registerQuartet<PA, PB, LC, LD, MODE_DENSE, 1>();
                }
            }
        }

        template <int LC>
        void regEri3()
        {
            if constexpr (!kF32Unit && kPairUnit < 10 && !(LC <= 3 && pairClassIdx(LC, 0) <= CI))
                registerQuartet<PA, PB, LC, 0, MODE_DENSE, 1>();
            // batched RI sums: eri3 (PA PB | LC), and eri2 (PA | LC) when PB = 0
            if constexpr (kPairUnit < 10)
                registerQuartet<PA, PB, LC, 0, MODE_SUM, kF32Unit ? 2 : 1>();
        }

        template <int LC>
        void regEri2()
        {
            if constexpr (!kF32Unit && kPairUnit >= 10 && LC <= PA)
            {
                registerQuartet<PA, 0, LC, 0, MODE_DENSE, 1>();
                registerQuartet<PA, 0, LC, 0, MODE_SUM, 1>();
            }
            if constexpr (kF32Unit && kPairUnit >= 10 && LC <= PA)
                registerQuartet<PA, 0, LC, 0, MODE_SUM, 2>();
        }

        struct Registrar
        {
            Registrar()
            {
                regEri4<0, 0>();
                regEri4<1, 0>();
                regEri4<1, 1>();
                regEri4<2, 0>();
                regEri4<2, 1>();
                regEri4<2, 2>();
                regEri4<3, 0>();
                regEri4<3, 1>();
                regEri4<3, 2>();
                regEri4<3, 3>();
                regEri3<0>();
                regEri3<1>();
                regEri3<2>();
                regEri3<3>();
                regEri3<4>();
                regEri3<5>();
                regEri2<0>();
                regEri2<1>();
                regEri2<2>();
                regEri2<3>();
                regEri2<4>();
                regEri2<5>();
                if constexpr (!kF32Unit)
                    registerQuartet<PA, PB, PA, PB, MODE_DIAG, 1>();
            }
        } registrar;
    }
}
