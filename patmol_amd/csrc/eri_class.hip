// eri_class.hip -- one angular-momentum class of the ERI kernels per object
// file.  Built once per (bra pair class, ket pair class) with
// -DPM_LA -DPM_LB -DPM_LC -DPM_LD so that the classes compile in parallel.
#include "eri_kernels_impl.h"

#define PM_CAT2(a, b, c, d, e) cls_##a##b##_##c##d##_##e
#define PM_NAME(la, lb, lc, ld, s) PM_CAT2(la, lb, lc, ld, s)

namespace pm {

int PM_NAME(PM_LA, PM_LB, PM_LC, PM_LD, jk)(const JKArgs &args, int prec, hipStream_t s)
{
    launch_jk_t<PM_LA, PM_LB, PM_LC, PM_LD>(args, prec, s);
    return 0;
}

int PM_NAME(PM_LA, PM_LB, PM_LC, PM_LD, out)(const EriOutArgs &args, hipStream_t s)
{
    launch_out_t<PM_LA, PM_LB, PM_LC, PM_LD>(args, s);
    return 0;
}

#if PM_LA == PM_LC && PM_LB == PM_LD
int PM_NAME(PM_LA, PM_LB, PM_LC, PM_LD, schwarz)(const SchwarzArgs &args, hipStream_t s)
{
    launch_schwarz_t<PM_LA, PM_LB>(args, s);
    return 0;
}
#endif

}  // namespace pm
