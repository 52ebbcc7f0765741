// mesh.hip -- triangle-mesh walls on the device.
//
// Particle-centric restatement of the reference's triangle-centric loops:
// each particle owns a row of candidate triangles (ascending id, the order in
// which FixWallGran::post_force_mesh visits a particle's triangles,
// fix_wall_gran.cpp:803-980) and as many contact slots
// (FixContactHistoryMesh, fix_contact_history_mesh_I.h:51-215).  One thread
// handles one particle, so the slot bookkeeping (first free slot, coplanar
// hand-over, clean-up) runs in the reference's order without atomics; G
// lanes of the particle share the geometry and the candidate search.
//
// k_mesh_nb<MODE,G> FixNeighlistMesh::handleTriangle (fix_neighlist_mesh.cpp:
//                   355-410) with resolveTriSphereNeighbuild, per particle:
//                   count, fill, or stage (count + first kMeshStage rows)
// k_mesh_compact    staged rows into the scanned list
// k_mesh_remap      FixContactHistoryMesh::sort_contacts (:375-402) and
//                   cleanUpContactJumps (:466-500) at a rebuild
// k_mesh_force<WN>  post_force_mesh + handleContact + cleanUpContacts
// k_mesh_move       MeshMoverLinear::initial_integrate (mesh_mover_linear.cpp:
//                   94-112)
#include "contact_law.h"
#include "kernels.h"
#include "mesh_geom.h"

namespace dem {

// G lanes per particle test the triangles of its grid cell G at a time; the
// passing ones are compacted by ballot + popcount, so a row keeps the cell
// list's (ascending id) order.  MODE kMCount counts the rows, kMFill writes
// them (after the scan), kMStage counts them and stages the first
// kMeshStage candidates of each, for k_mesh_compact.
constexpr int kMCount = 0, kMFill = 1, kMStage = 2;

template <int MODE, int G>
__global__ void __launch_bounds__(kBlock) k_mesh_nb(MeshNbArgs a)
{
    constexpr bool FILL = MODE == kMFill;
    constexpr bool STAGE = MODE == kMStage;
    const int lane = threadIdx.x & (G - 1);
    const int gshift = (threadIdx.x & 63) & ~(G - 1);
    const int i = (blockIdx.x * kBlock + threadIdx.x) / G;
    if (i >= a.n) return;   // uniform across the group
    // recorded after a staged pass: fills only when a row outgrew its stage
    if (FILL && a.fill_if_over && a.mstart[i + 1] - a.mstart[i] <= kMeshStage) return;
    const double4 p = a.xr[i];
    const double q[3] = {p.x - a.disp[0], p.y - a.disp[1], p.z - a.disp[2]};
    int c[3];
    bool in = true;
    for (int d = 0; d < 3; d++) {
        const double f = floor((q[d] - a.glo[d]) * a.ginv[d]);
        in = in && f >= 0.0 && f < (double)a.gn[d];
        c[d] = in ? (int)f : 0;
    }
    int k = 0;
    if (in) {
        const int cell = (c[2] * a.gn[1] + c[1]) * a.gn[0] + c[0];
        const double cs[3] = {p.x, p.y, p.z};
        const double r = p.w * a.cdf;
        const int base = FILL ? a.mstart[i] : 0;
        const int e0 = a.gstart[cell], e1 = a.gstart[cell + 1];
        for (int eb = e0; eb < e1; eb += G) {
            const int e = eb + lane;
            int t = 0;
            bool pass = false;
            if (e < e1) {
                t = a.glist[e];
                pass = tri_neighbuild(a.tri[t], a.enorm + 9 * (size_t)t, r, cs, a.threshold);
            }
            const uint64_t m = (__ballot(pass) >> gshift) & ((1ull << G) - 1ull);
            const int pos = k + __popcll(m & ((1ull << lane) - 1ull));
            if (FILL && pass) a.mcand[base + pos] = t;
            if (STAGE && pass && pos < kMeshStage) a.stage[(size_t)i * kMeshStage + pos] = t;
            k += __popcll(m);
        }
    }
    if (!FILL && lane == 0) a.count[i] = k;
}

// The staged rows into mcand once the row lengths are scanned, G lanes per
// particle.  Nothing when a row outgrew its stage (the fill pass runs then).
template <int G>
__global__ void __launch_bounds__(kBlock) k_mesh_compact(MeshNbArgs a)
{
    const int lane = threadIdx.x & (G - 1);
    const int i = (blockIdx.x * kBlock + threadIdx.x) / G;
    if (i >= a.n) return;
    const int b0 = a.mstart[i];
    const int cnt = a.mstart[i + 1] - b0;
    if (cnt > kMeshStage) return;   // re-searched by the fill pass
    for (int k = lane; k < cnt; k += G) a.mcand[b0 + k] = a.stage[(size_t)i * kMeshStage + k];
}

__global__ void __launch_bounds__(kBlock) k_mesh_remap(MeshRemapArgs a)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const int base = a.mstart[i];
    const int nn = a.mstart[i + 1] - base;
    const int dnum = a.dnum;
    int np = 0, onn = 0;
    int *pt = nullptr;
    double *oh = nullptr;
    if (a.have_old) {
        const int o = a.perm[i];
        const int ob = a.ostart[o];
        onn = a.ostart[o + 1] - ob;
        pt = a.opart + ob;
        oh = a.ohist + (size_t)ob * dnum;
        np = a.onp[o];
        // sort_contacts: first empty <-> last filled until compact
        for (;;) {
            int fe = -1, lf = -1;
            for (int j = 0; j < onn; j++) {
                if (fe == -1 && pt[j] == -1) fe = j;
                if (pt[j] >= 0) lf = j;
            }
            if (!(fe > -1 && lf > -1 && fe < lf)) break;
            const int tt = pt[fe];
            pt[fe] = pt[lf];
            pt[lf] = tt;
            for (int d = 0; d < dnum; d++) {
                const double x = oh[(size_t)fe * dnum + d];
                oh[(size_t)fe * dnum + d] = oh[(size_t)lf * dnum + d];
                oh[(size_t)lf * dnum + d] = x;
            }
        }
        // cleanUpContactJumps: a partner that is no longer a candidate is
        // cleared and swapped with the last filled slot
        int ip = 0;
        while (ip < np && ip < onn) {
            const int tri = pt[ip];
            bool found = false;
            for (int k = 0; k < nn; k++)
                if (a.mcand[base + k] == tri) {
                    found = true;
                    break;
                }
            if (found) {
                ip++;
                continue;
            }
            pt[ip] = -1;
            for (int d = 0; d < dnum; d++) oh[(size_t)ip * dnum + d] = 0.0;
            const int last = np - 1;
            const int tt = pt[ip];
            pt[ip] = pt[last];
            pt[last] = tt;
            for (int d = 0; d < dnum; d++) {
                const double x = oh[(size_t)ip * dnum + d];
                oh[(size_t)ip * dnum + d] = oh[(size_t)last * dnum + d];
                oh[(size_t)last * dnum + d] = x;
            }
            np--;
        }
    }
    for (int k = 0; k < nn; k++) {
        const bool carry = k < np && k < onn;
        a.mpart[base + k] = carry ? pt[k] : -1;
        for (int d = 0; d < dnum; d++)
            a.mhist[(size_t)(base + k) * dnum + d] = carry ? oh[(size_t)k * dnum + d] : 0.0;
    }
    a.mnp[i] = np;
}

// During a force pass a slot's keep flag (FixContactHistoryMesh keepflag_)
// rides in bit 30 of its partner id; -1 marks an empty slot.
constexpr int kKeep = 1 << 30;
constexpr int kIdMask = kKeep - 1;

// G lanes per particle evaluate the geometry of its candidate triangles G
// at a time (no stores, so the gathers of a whole group are in flight
// together); the candidates within reach then go, in ascending order, to
// lane 0, which keeps the contact slots and evaluates the contacts exactly
// as the one-thread form did.
template <int WN, int G>
__global__ void __launch_bounds__(kBlock) k_mesh_force(MeshForceArgs a)
{
    const int lane = threadIdx.x & (G - 1);
    const int gshift = (threadIdx.x & 63) & ~(G - 1);
    const int i = (blockIdx.x * kBlock + threadIdx.x) / G;
    if (i >= a.n) return;   // uniform across the group
    if (a.halt && *a.halt) return;
    const Model &M = a.M;
    const int base = a.mstart[i];
    const int nn = a.mstart[i + 1] - base;
    const double4 wi = a.wt[i];
    const int meta = (int)wi.w;
    const bool frozen = a.apply_freeze && M.freeze_on && meta_frozen(meta);
    // fix mesh/surface/stress sums of this particle's contacts
    double SF[3] = {0.0, 0.0, 0.0}, ST[3] = {0.0, 0.0, 0.0};
    if (nn > 0) {
        const double4 pi = a.xr[i];
        const double4 vmi = a.vm[i];
        const double cs[3] = {pi.x, pi.y, pi.z};
        const double radi = pi.w;
        const double cdm = M.cdf - 1.0;
        const int dnum = M.wdnum;
        int *part = a.mpart + base;
        double *hist = a.mhist + (size_t)base * dnum;
        int np = a.mnp[i];
        double F[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0};
        // FixContactHistoryMesh::handleContact and the contact, lane 0
        auto handle = [&](int t, double deltan, const double *delta, const double *bary, int region) {
            const TriGeo &g = a.tri[t];
            const bool intersect = deltan <= 0;
            int slot = -1;
            for (int s = 0; s < nn; s++) {
                const int pv = part[s];
                if (pv >= 0 && (pv & kIdMask) == t) {
                    slot = s;
                    break;
                }
            }
            if (slot >= 0) {
                part[slot] = t | kKeep;
            } else {
                const bool face = region == 7;
                if (face) {
                    bool already = false;
                    for (int s = 0; s < nn && !already; s++) {
                        const int pv = part[s];
                        if (pv < 0 || !(pv & kKeep)) continue;
                        const int pp = pv & kIdMask;
                        if (pp != t && tri_coplanar_node_neighs(a.tri[pp], t, g, a.curvature, a.precision))
                            already = true;
                    }
                    if (already) return;
                }
                for (int s = 0; s < nn; s++)
                    if (part[s] == -1) {
                        slot = s;
                        break;
                    }
                if (slot < 0 || np >= nn) {
                    atomicOr(a.err, 4);
                    return;
                }
                part[slot] = t | kKeep;
                double *h = hist + (size_t)slot * dnum;
                for (int d = 0; d < dnum; d++) h[d] = 0.0;
                np++;
                if (face)
                    for (int s = 0; s < nn; s++) {
                        const int pv = part[s];
                        if (pv < 0) continue;
                        const int pp = pv & kIdMask;
                        if (pp != t && tri_coplanar_node_neighs(a.tri[pp], t, g, a.curvature, a.precision))
                            for (int d = 0; d < dnum; d++) h[d] = hist[(size_t)s * dnum + d];
                    }
            }
            double *hp = hist + (size_t)slot * dnum;
            if (intersect) {
                ContactIn cw;
                cw.is_wall = true;
                cw.itype = meta_type(meta);
                cw.jtype = a.wtype;
                cw.radi = radi;
                cw.radj = 0.0;
                cw.deltan_in = -deltan;
                cw.delta[0] = -delta[0]; cw.delta[1] = -delta[1]; cw.delta[2] = -delta[2];
                cw.r = cw.radi - cw.deltan_in;
                cw.rinv = 1.0 / cw.r;
                cw.radsum = cw.radi;
                cw.mi = vmi.w; cw.mj = 0.0; cw.meff = vmi.w;
                cw.vi[0] = vmi.x; cw.vi[1] = vmi.y; cw.vi[2] = vmi.z;
                for (int q = 0; q < 3; q++)
                    cw.vj[q] = a.moving ? (bary[0] * a.vel[q] + bary[1] * a.vel[q] + bary[2] * a.vel[q]) : 0.0;
                cw.wi[0] = wi.x; cw.wi[1] = wi.y; cw.wi[2] = wi.z;
                cw.wj[0] = 0.0; cw.wj[1] = 0.0; cw.wj[2] = 0.0;
                constexpr int WHN = chain_max_dnum(WN);
                double h[WHN];
#pragma unroll
                for (int q = 0; q < WHN; q++) h[q] = q < dnum ? hp[q] : 0.0;
                double dF[3], dT[3];
                contact_fp64<WN>(M, M.lw, a.mat, cw, h, a.shearupdate, dF, dT);
                if (a.shearupdate || WN != NM_HERTZ) {
#pragma unroll
                    for (int q = 0; q < WHN; q++)
                        if (q < dnum) hp[q] = h[q];
                }
                F[0] += dF[0]; F[1] += dF[1]; F[2] += dF[2];
                T[0] += dT[0]; T[1] += dT[1]; T[2] += dT[2];
                if (a.mstress) {
                    // MeshModuleStress::add_particle_contribution
                    // (mesh_module_stress.cpp:318-344): -F at the contact
                    // point x + delta, torque about p_ref
                    const double frc[3] = {-dF[0], -dF[1], -dF[2]};
                    const double arm[3] = {cs[0] + delta[0] - a.pref[0], cs[1] + delta[1] - a.pref[1],
                                           cs[2] + delta[2] - a.pref[2]};
                    SF[0] += frc[0]; SF[1] += frc[1]; SF[2] += frc[2];
                    ST[0] += arm[1] * frc[2] - arm[2] * frc[1];
                    ST[1] += arm[2] * frc[0] - arm[0] * frc[2];
                    ST[2] += arm[0] * frc[1] - arm[1] * frc[0];
                }
            } else {
                contact_close(M.lw, hp);
            }
        };
        for (int k0 = 0; k0 < nn; k0 += G) {
            const int k = k0 + lane;
            double delta[3] = {0.0, 0.0, 0.0}, bary[3] = {0.0, 0.0, 0.0};
            double deltan = 0.0;
            int region = -1, t = 0;
            bool reach = false;
            if (k < nn) {
                t = a.mcand[base + k];
                deltan = tri_contact(a.tri[t], radi, cs, a.precision, delta, bary, region);
                reach = !(deltan > M.cutneighmax) && (deltan <= 0 || deltan < cdm * radi);
            }
            uint64_t m = (__ballot(reach) >> gshift) & ((1ull << G) - 1ull);
            while (m) {   // the same mask on every lane of the group
                const int b = __ffsll((long long)m) - 1;
                m &= m - 1;
                // every lane of the group takes part in the shuffles
                const int tb = __shfl(t, b, G);
                const double db = __shfl(deltan, b, G);
                const int rb = __shfl(region, b, G);
                double dl[3], br[3];
                for (int q = 0; q < 3; q++) {
                    dl[q] = __shfl(delta[q], b, G);
                    br[q] = __shfl(bary[q], b, G);
                }
                if (lane == 0) handle(tb, db, dl, br, rb);
            }
        }
        if (lane == 0) {
            // cleanUpContacts; a slot that is already empty holds zero
            // history (it was zeroed when it emptied), so it is left alone
            for (int s = 0; s < nn; s++) {
                const int pv = part[s];
                if (pv < 0) continue;
                if (pv & kKeep) {
                    part[s] = pv & kIdMask;
                    continue;
                }
                np--;
                part[s] = -1;
                for (int d = 0; d < dnum; d++) hist[(size_t)s * dnum + d] = 0.0;
            }
            a.mnp[i] = np;
            if (!frozen) {
                double4 f = a.f4[i], tq = a.t4[i];
                f.x += F[0]; f.y += F[1]; f.z += F[2];
                tq.x += T[0]; tq.y += T[1]; tq.z += T[2];
                a.f4[i] = f;
                a.t4[i] = tq;
            }
        }
    }
    if (frozen && lane == 0) {
        a.f4[i] = make_double4(0.0, 0.0, 0.0, 0.0);
        a.t4[i] = make_double4(0.0, 0.0, 0.0, 0.0);
    }
    if (a.mstress && lane == 0) {
        double *o = a.mstress + 6 * (size_t)i;
        o[0] = SF[0]; o[1] = SF[1]; o[2] = SF[2];
        o[3] = ST[0]; o[4] = ST[1]; o[5] = ST[2];
    }
    // the last force term of the step: the integrator's operations
    // (k_integrate) on the final force and torque, into the other buffers
    if (a.fuse && lane == 0) {
        const double4 f = a.f4[i], t = a.t4[i];
        double4 x = a.xr[i], v = a.vm[i], w = wi;
        // dtfm = dtf / (rmass*(1.+CAddRhoFluid/density)), CAddRhoFluid = 0
        const double dtfm = a.dtf / (v.w * (1. + 0.0 / 1.0));
        const double dtirotate = a.dtfrotate / (x.w * x.w * v.w);
        if (a.fuse & 2) {
            v.x += dtfm * f.x; v.y += dtfm * f.y; v.z += dtfm * f.z;
            w.x += dtirotate * t.x; w.y += dtirotate * t.y; w.z += dtirotate * t.z;
        }
        if (a.fuse & 1) {
            v.x += dtfm * f.x; v.y += dtfm * f.y; v.z += dtfm * f.z;
            x.x += a.dtv * v.x; x.y += a.dtv * v.y; x.z += a.dtv * v.z;
            w.x += dtirotate * t.x; w.y += dtirotate * t.y; w.z += dtirotate * t.z;
        }
        a.xn[i] = x;
        a.vn[i] = v;
        a.wn[i] = w;
    }
}

__global__ void __launch_bounds__(kBlock) k_mesh_move(TriGeo *tri, int ntri, double dx, double dy, double dz,
                                                    const long long *halt)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= ntri) return;
    if (halt && *halt) return;
    move_tri(tri[t], dx, dy, dz);
}

static int blocks(long long n) { return (int)((n + kBlock - 1) / kBlock); }

constexpr int kMeshLanes = 8;   // lanes per particle in the mesh kernels

void launch_mesh_count(const MeshNbArgs &a, hipStream_t s)
{
    if (a.n) k_mesh_nb<kMCount, kMeshLanes><<<blocks((long long)a.n * kMeshLanes), kBlock, 0, s>>>(a);
}

void launch_mesh_fill(const MeshNbArgs &a, hipStream_t s)
{
    if (a.n) k_mesh_nb<kMFill, kMeshLanes><<<blocks((long long)a.n * kMeshLanes), kBlock, 0, s>>>(a);
}

void launch_mesh_stage(const MeshNbArgs &a, hipStream_t s)
{
    if (a.n) k_mesh_nb<kMStage, kMeshLanes><<<blocks((long long)a.n * kMeshLanes), kBlock, 0, s>>>(a);
}

void launch_mesh_compact(const MeshNbArgs &a, hipStream_t s)
{
    if (a.n) k_mesh_compact<kMeshLanes><<<blocks((long long)a.n * kMeshLanes), kBlock, 0, s>>>(a);
}

void launch_mesh_remap(const MeshRemapArgs &a, hipStream_t s)
{
    if (a.n) k_mesh_remap<<<blocks(a.n), kBlock, 0, s>>>(a);
}

void launch_mesh_force(const MeshForceArgs &a, hipStream_t s)
{
    if (!a.n) return;
    const int nb = blocks((long long)a.n * kMeshLanes);
    switch (a.M.lw.normal) {
    case NM_HYST1: k_mesh_force<NM_HYST1, kMeshLanes><<<nb, kBlock, 0, s>>>(a); break;
    case NM_HYST2: k_mesh_force<NM_HYST2, kMeshLanes><<<nb, kBlock, 0, s>>>(a); break;
    default: k_mesh_force<NM_HERTZ, kMeshLanes><<<nb, kBlock, 0, s>>>(a); break;
    }
}

void launch_mesh_move(TriGeo *tri, int ntri, double dx, double dy, double dz, hipStream_t s, const long long *halt)
{
    if (ntri) k_mesh_move<<<blocks(ntri), kBlock, 0, s>>>(tri, ntri, dx, dy, dz, halt);
}

// Fixed-order sum of a per-particle 6-vector: each of kSumBlocks blocks
// sums a contiguous range (a fixed tree inside the block), one block adds
// the partial sums in order.  Deterministic; no atomics.
__global__ void __launch_bounds__(kBlock) k_sum6_part(const double *v, int n, double *part)
{
    __shared__ double s_v[6][kBlock];
    const int per = (n + gridDim.x - 1) / gridDim.x;
    const int p0 = blockIdx.x * per, p1 = min(n, p0 + per);
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int p = p0 + (int)threadIdx.x; p < p1; p += kBlock)
        for (int k = 0; k < 6; k++) acc[k] += v[6 * (size_t)p + k];
    for (int k = 0; k < 6; k++) s_v[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int k = 0; k < 6; k++) s_v[k][threadIdx.x] += s_v[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s_v[threadIdx.x][0];
}

__global__ void k_sum6_total(const double *part, int nb, double *out)
{
    if (threadIdx.x >= 6) return;
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += part[6 * b + threadIdx.x];
    out[threadIdx.x] = s;
}

void launch_sum6(const double *v, int n, double *part, double *out, hipStream_t s)
{
    k_sum6_part<<<kSumBlocks, kBlock, 0, s>>>(v, n, part);
    k_sum6_total<<<1, 64, 0, s>>>(part, kSumBlocks, out);
}

// FixMassflowMesh::post_integrate for a planar face (no point_at_outlet):
// the side of a particle is the sign of (x - p_ref) . n, n the face normal
// turned towards vec_side.  A particle in some triangle's neighbour list
// that is seen for the first time only records its side; one that has
// moved to the n side and lies in the face region (barySign 7) of one of
// its triangles counts once per crossing (or once ever with `count once`);
// one on the other side is re-armed.  The triangles of a particle are
// independent here, so one pass per particle over its candidate row gives
// the reference's triangle-by-triangle result.  The step's mass and count
// are summed in a fixed order and added to the running totals.
__global__ void __launch_bounds__(kBlock) k_massflow(MassflowArgs a)
{
    __shared__ double s_m[kBlock], s_c[kBlock];
    double mass = 0.0, cnt = 0.0;
    const int per = (a.n + gridDim.x - 1) / gridDim.x;
    const int p0 = blockIdx.x * per, p1 = min(a.n, p0 + per);
    for (int i = p0 + (int)threadIdx.x; i < p1; i += kBlock) {
        if (a.gmask && !((a.gmask[i] >> a.gbit) & 1)) continue;
        const int c = a.counter[i];
        if (c == 2) continue;
        const int b0 = a.mstart[i], b1 = a.mstart[i + 1];
        if (b1 == b0) continue;   // in no triangle's neighbour list
        const double4 x = a.xr[i];
        const double dot = (x.x - a.pref[0]) * a.nvec[0] + (x.y - a.pref[1]) * a.nvec[1] + (x.z - a.pref[2]) * a.nvec[2];
        if (c == -1) {
            a.counter[i] = dot <= 0.0 ? 0 : 1;
            continue;
        }
        if (dot > 0.0) {
            if (c != 0) continue;
            const double cs[3] = {x.x, x.y, x.z};
            bool face = false;
            for (int k = b0; k < b1 && !face; k++) {
                double delta[3], bary[3];
                int region = -1;
                tri_contact(a.tri[a.mcand[k]], x.w, cs, a.precision, delta, bary, region);
                face = region == 7;
            }
            if (face) {
                mass += a.vm[i].w;
                cnt += 1.0;
                a.counter[i] = a.once ? 2 : 1;
            }
        } else {
            a.counter[i] = 0;
        }
    }
    s_m[threadIdx.x] = mass;
    s_c[threadIdx.x] = cnt;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s_m[threadIdx.x] += s_m[threadIdx.x + h];
            s_c[threadIdx.x] += s_c[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.part[2 * blockIdx.x] = s_m[0];
        a.part[2 * blockIdx.x + 1] = s_c[0];
    }
}

__global__ void k_massflow_total(const double *part, int nb, double *total)
{
    if (threadIdx.x >= 2) return;
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += part[2 * b + threadIdx.x];
    total[threadIdx.x] += s;
}

void launch_massflow(const MassflowArgs &a, hipStream_t s)
{
    if (a.n == 0) return;
    k_massflow<<<kMassflowBlocks, kBlock, 0, s>>>(a);
    k_massflow_total<<<1, 64, 0, s>>>(a.part, kMassflowBlocks, a.total);
}

}  // namespace dem
