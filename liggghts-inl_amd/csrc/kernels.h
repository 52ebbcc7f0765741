// kernels.h -- host-callable launchers of the engine's HIP kernels.
#pragma once
#include "dem_device.h"

namespace dem {

// Everything the force pass reads and writes (one particle per thread).
struct ForceArgs {
    const double4 *xr, *vm, *wt;
    const int *nstart;
    uint32_t *nj;
    double *hist;
    double4 *f4, *t4;
    const double *mat;
    const uint8_t *wmask;
    double *whist;
    int *ncontact;       // optional per-atom contact count (stats), may be null
    int *err;            // device error flags
    int n;
    int shearupdate;
    Model M;
    Planes P;
};

struct BinArgs {
    double4 *xr;         // wrapped in place (Domain::pbc)
    int4 *img;
    uint32_t *key;
    int *count;          // per Morton key
    int n;
    Grid G;
    double prd[3];
    double hi[3];
};

struct PermuteArgs {
    const int *sorted;   // new -> old
    const double4 *xr0, *vm0, *wt0;
    const int *tag0, *orig0;
    const int4 *img0;
    const double *whist0;
    double4 *xr1, *vm1, *wt1, *xhold;
    int *tag1, *orig1;
    int4 *img1;
    double *whist1;
    uint8_t *wmask;
    int n;
    int wdn;             // nplanes * dnum
    Planes P;
    double skin;
};

struct NeighArgs {
    const double4 *xr;   // new order
    const int *tag;      // new order
    const int *cstart;   // per key, exclusive prefix + total at nkeys
    int *count;          // per atom (count pass)
    int *maxcount;
    const int *nstart;   // per atom (fill pass)
    uint32_t *nj;
    double *hist;
    // previous list, old index space
    const int *perm;     // new -> old
    const int *onstart;
    const uint32_t *onj;
    const double *ohist;
    const int *otag;
    int have_old;
    int n;
    int dnum;
    double cdf, skin;
    Grid G;
    double prd[3];
};

struct IntegArgs {
    double4 *xr, *vm, *wt;
    const double4 *f4, *t4;
    int n;
    double dtf, dtv, dtfrotate;
};

void launch_bin(const BinArgs &a, hipStream_t s);
void launch_scan(int *data, int n, int *tmp, hipStream_t s);   // exclusive, data[n] = total
void launch_scatter(const uint32_t *key, int *cursor, int *sorted, int n, hipStream_t s);
void launch_cellsort(const int *cstart, int *sorted, int nkeys, hipStream_t s);
void launch_permute(const PermuteArgs &a, hipStream_t s);
void launch_neigh_count(const NeighArgs &a, hipStream_t s);
void launch_neigh_fill(const NeighArgs &a, hipStream_t s);
void launch_force_fp64(const ForceArgs &a, hipStream_t s);
void launch_force_mixed(const ForceArgs &a, hipStream_t s);
// mode: 1 = initial_integrate, 2 = final_integrate, 3 = final then initial
void launch_integrate(const IntegArgs &a, int mode, hipStream_t s);
int scan_tmp_size(int n);
// Neighbor::check_distance (neighbor.cpp:1390-1465): flag |= 1 if any atom
// moved more than sqrt(deltasq) since the last build.
void launch_check_distance(const double4 *xr, const double4 *xhold, int n, double deltasq, int *flag,
                           hipStream_t s);

}  // namespace dem
