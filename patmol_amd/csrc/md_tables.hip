// md_tables.hip -- launcher of the Hermite-coefficient table kernels of the
// McMurchie-Davidson J/K path (md_etab_kernel, md_kernel.h) for the pair
// classes it reads: dd, dp (bra and ket tables) and pp (ket table).
#include "eri_kernels_impl.h"

namespace pm {

int launch_md_etab(int la, int lb, const MdTabArgs &args, hipStream_t s)
{
    if (args.npairs <= 0) return 0;
    const dim3 grid((unsigned)args.npairs), block(64);
    if (la == 2 && lb == 2) hipLaunchKernelGGL((md_etab_kernel<2, 2>), grid, block, 0, s, args);
    else if (la == 2 && lb == 1) hipLaunchKernelGGL((md_etab_kernel<2, 1>), grid, block, 0, s, args);
    else if (la == 1 && lb == 1) hipLaunchKernelGGL((md_etab_kernel<1, 1>), grid, block, 0, s, args);
    else return 1;
    return 0;
}

}  // namespace pm
