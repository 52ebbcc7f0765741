// dem_device.h -- device-side data layout shared by the engine's kernels.
//
// HBM layout (all arrays in Morton cell order after each rebuild):
//   xr  double4  (x, y, z, radius)           32 B / particle
//   vm  double4  (vx, vy, vz, rmass)         32 B
//   wt  double4  (wx, wy, wz, meta)          32 B   meta = type | freeze<<4 | instance<<5
//   f4  double4  (fx, fy, fz, -)             32 B
//   t4  double4  (tx, ty, tz, -)             32 B
//   tag int, orig int (upload position), img int4 (periodic image counts)
// Neighbour list: full (both directions), CSR rows per particle:
//   nstart[n+1]; nj[e] = j | bond << 25 | shift << 26 | flag << 31; hist[e*dnum + k]
//   (history slots: hysteretic 12, shear 3, rolling 3, cohesion 14/28)
// Primitive walls: wmask[i] bit k = atom in wall k's list; whist[(i*np+k)*dnum]
// Mesh walls, per mesh: TriGeo records; per particle a CSR row (mstart) of
// candidate triangles (mcand, ascending id) and as many contact slots
// (mpart = triangle id or -1, mhist = wall history), mnp = filled slots
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dem {

constexpr int kMaxTypes = 8;
constexpr int kT1 = kMaxTypes + 1;
constexpr int kMaxPlanes = 8;
constexpr int kMaxDnum = 18;   // hysteretic 12 + shear 3 + rolling 3 (walls)
constexpr uint32_t kJMask = (1u << 25) - 1u;
constexpr uint32_t kBondBit = 1u << 25;   // the cohesion model's contact flag
constexpr uint32_t kFlagBit = 1u << 31;   // the other models' contact flags
constexpr int kNoShift = 13;
constexpr int kBlock = 256;

// material table slots, each kT1*kT1 doubles, index [itype*kT1 + jtype];
// the hysteretic/nonlinear* parameters follow in the order of
// normal_model_hysteretic_nonlinear1.h:105-133
enum {
    M_YEFF = 0, M_GEFF, M_BETA, M_MU, M_MUR, M_ETAR, M_RESTLOG,
    M_HK, M_HKN2K1, M_HKN2KC, M_HPHIF, M_HFADH, M_HALPHA, M_HCIN, M_HA1, M_HA2, M_HA3, M_HKCIN,
    // bond tables (cohesion bond, bond/nonlinear), in the order of DEM_NBOND
    M_BLAMBDA, M_BKN, M_BKT, M_BDFN, M_BDFT, M_BDTN, M_BDTT, M_BMAXDIST, M_BMAXSIGMA, M_BMAXTAU,
    M_BRATIOTC, M_BCREATEDIST, M_BKFN1, M_BKUFN1, M_BKCFN1, M_BKFN2, M_BKUFN2, M_BKCFN2, M_BKFT,
    M_BKTN, M_BKUTN, M_BKCTN, M_BKTT, M_BKUTT, M_BKCTT,
    M_COUNT
};
constexpr int kNHyst = 11;
constexpr int kNBond = 25;
// one instance of a batched run owns one block of material tables
constexpr int kMatStride = M_COUNT * kT1 * kT1;

// wt.w packs the atom type (bits 0-3), fix freeze membership (bit 4) and the
// instance of a batched run (bits 5 and up; 0 for a single run)
__host__ __device__ inline int meta_type(int m) { return m & 15; }
__host__ __device__ inline bool meta_frozen(int m) { return ((m >> 4) & 1) != 0; }
__host__ __device__ inline int meta_inst(int m) { return m >> 5; }

// cohesion model of the pair chain
enum { CM_OFF = 0, CM_BOND = 1, CM_BOND_NONLINEAR = 2 };

// normal model of a contact chain
enum { NM_HERTZ = 0, NM_HYST1 = 1, NM_HYST2 = 2 };

// history layout of a contact chain (slots in registration order:
// normal model, tangential history, rolling)
struct Layout {
    int normal;      // NM_*
    int limitf;      // limitForce
    int off_norm;    // hysteretic: 12 slots, else -1
    int off_shear;
    int off_roll;    // -1 without rolling friction
    int off_coh;     // cohesion slots (after the others on the device), -1 none
    int dnum;
};

// bond settings (registerSettings of the cohesion models)
struct BondSet {
    long long ts_create;   // static_cast<int>(tsCreateBond)
    int stress_break, create_always, damping, damping_smooth, ratio_tc, tension_variant;
};

struct Planes {
    int n;
    int dim[kMaxPlanes];
    int wtype[kMaxPlanes];
    double pos[kMaxPlanes];
};

// Triangle of a mesh wall, one 256-byte record per triangle (the contact
// pass reads one record per candidate).  flags: bit k edge k active, bit
// 3+k corner k active.  neigh: edge neighbours (first five, mesh ids).
constexpr int kMaxMeshes = 4;
constexpr int kMaxMeshSlots = 4096; // candidate triangles (= contact slots) per particle and mesh
constexpr int kMeshNeigh = 5;
struct TriGeo {
    double node[9];
    double evec[9];
    double elen[3];
    double snorm[3];
    double center[3];
    double rbound;
    int obtuse;
    int flags;
    int nneigh;
    int neigh[kMeshNeigh];
};
static_assert(sizeof(TriGeo) == 256, "TriGeo is one 256-byte record");

// MeshMoverLinear's step of one triangle: every node and the centre moved by
// (dx, dy, dz), one addition per component
__device__ inline void move_tri(TriGeo &g, double dx, double dy, double dz)
{
    for (int k = 0; k < 3; k++) {
        g.node[3 * k] = g.node[3 * k] + dx;
        g.node[3 * k + 1] = g.node[3 * k + 1] + dy;
        g.node[3 * k + 2] = g.node[3 * k + 2] + dz;
    }
    g.center[0] = g.center[0] + dx;
    g.center[1] = g.center[1] + dy;
    g.center[2] = g.center[2] + dz;
}

struct Model {
    double dt;
    double cdf;          // contactDistanceFactor
    double skin;
    double cutneighmax;
    double prd[3];
    double gacc[3];
    double krs;          // coefficientRollingStiffness (epsd3)
    int gravity_on;
    int freeze_on;
    int rolling;         // 0 off, 1 epsd, 2 epsd2, 3 epsd3, 4 cdtnonlinear2
    int tdamp;           // tangential_damping
    int torsion;         // rolling torsionTorque
    int dnum;            // history doubles per pair
    int wdnum;           // history doubles per (atom, plane)
    Layout lp, lw;       // pair and wall chains
    int cohesion;        // CM_* of the pair chain
    BondSet bs;
};

struct Grid {
    double lo[3];
    double inv[3];       // cells per unit length
    int nc[3];
    int periodic[3];
    int nkeys;           // size of the key space (all instances)
    int nkeys_local;     // Morton keys of one instance; key = instance*nkeys_local + morton
    const uint32_t *spread;  // spread[d*maxc + c] = interleaved bits of cell c in dim d
    int maxc;
    int reach;           // cells per cutneighmax: the stencil is 2*reach+1 cells wide
};

__host__ __device__ inline void shift_decode(int code, int &sx, int &sy, int &sz)
{
    sx = code % 3 - 1;
    sy = (code / 3) % 3 - 1;
    sz = code / 9 - 1;
}

// x[j] + dx with dx = pbc * prd, the expression of comm->borders()
// (atom_vec_sphere.cpp:625-637)
__device__ inline double3 image_of(const double4 &p, int code, const double prd[3])
{
    int sx, sy, sz;
    shift_decode(code, sx, sy, sz);
    double3 r;
    r.x = p.x + sx * prd[0];
    r.y = p.y + sy * prd[1];
    r.z = p.z + sz * prd[2];
    return r;
}

}  // namespace dem
