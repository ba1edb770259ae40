// et_sweep_r.hip -- instantiates the sweep launchers of one PSF radius.
// Compiled once per radius 0..ET_RMAX with -DET_SWEEP_R=<radius>.
#include "et_sweep.h"

#ifndef ET_SWEEP_R
#error "compile with -DET_SWEEP_R=<PSF radius>"
#endif
#define ET_SWEEP_CAT2(a, b) a##b
#define ET_SWEEP_CAT(a, b) ET_SWEEP_CAT2(a, b)

int ET_SWEEP_CAT(et_sweep_launch_r, ET_SWEEP_R)(int mode, int has_mu, const EtSweepArgs &a, hipStream_t st,
                                                const GroupLaunch *grp)
{
    return launch_sweep_r<ET_SWEEP_R>(mode, has_mu, a, st, grp);
}

int ET_SWEEP_CAT(et_yl_launch_r, ET_SWEEP_R)(int kind, int RPL, int NR, int nthr, size_t lds, const EtSweepArgs *args,
                                              const int4 *tab, int n, int blocks, hipStream_t st)
{
    return launch_yl_r<ET_SWEEP_R>(kind, RPL, NR, nthr, lds, args, tab, n, blocks, st);
}
