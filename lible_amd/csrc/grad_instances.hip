// Instantiation unit of the gradient kernels, compiled once per LBX_GUNIT = lane-side main pair
// class (order (0,0),(1,0),(1,1),(2,0),...): one kernel for every uniform-side class for which
// this class is the lane side (gradLaneFirst).
#include "grad.hpp"

#ifndef LBX_GUNIT
#error "LBX_GUNIT must be defined"
#endif

namespace lbx
{
    namespace
    {
        constexpr int unitLa()
        {
            int la = 0;
            while (pairClassIdx(la + 1, 0) <= LBX_GUNIT)
                la++;
            return la;
        }
        constexpr int PA = unitLa(), PB = LBX_GUNIT - pairClassIdx(PA, 0);

        template <int LC, int LD>
        void reg()
        {
            if constexpr (gradLaneFirst(PA, PB, LC, LD))
                registerGrad<PA, PB, LC, LD>();
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
