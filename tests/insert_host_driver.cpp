// CPU driver of the engine's insertion code (liggghts-inl_amd/csrc/insert_pack.cpp)
// for tests/test_insert.py and tests/test_template.py: the device hooks
// replaced by brute force, so the host logic (batched attempts, stream
// rewinds, the serial walk; the template's bounding sphere, centre of mass
// and placement) can be checked against the oracle without a GPU.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../liggghts-inl_amd/csrc/insert_pack.h"

namespace {
struct Run {
    std::unique_ptr<dem_ins::InsertPack> fix;
    double lo[3], hi[3];
    int periodic[3];
};

// TriMesh::isInElement and FixInsertStream::is_nearby with the load_xnear
// bounding box, as the device filter applies them
bool in_element(const double *nd, const double *p)
{
    double v0[3], v1[3], v2[3];
    for (int q = 0; q < 3; q++) {
        v0[q] = nd[6 + q] - nd[q];
        v1[q] = nd[3 + q] - nd[q];
        v2[q] = p[q] - nd[q];
    }
    const double d00 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2];
    const double d01 = v0[0] * v1[0] + v0[1] * v1[1] + v0[2] * v1[2];
    const double d02 = v0[0] * v2[0] + v0[1] * v2[1] + v0[2] * v2[2];
    const double d11 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
    const double d12 = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
    const double inv = 1. / (d00 * d11 - d01 * d01);
    const double u = (d11 * d02 - d01 * d12) * inv;
    const double v = (d00 * d12 - d01 * d02) * inv;
    return u > -1e-10 && v > -1e-10 && u + v < 1. + 1e-10;
}

bool nearby(const dem_ins::FaceFilter &f, const double *x)
{
    double r[3];
    for (int q = 0; q < 3; q++) r[q] = x[q] - f.pref[q];
    const double d = r[0] * f.nrm[0] + r[1] * f.nrm[1] + r[2] * f.nrm[2];
    if (d > f.dmax || d < f.dmin) return false;
    double pr[3];
    for (int q = 0; q < 3; q++) pr[q] = x[q] + f.nrm[q] * d;
    bool on = false;
    for (size_t t = 0; t < f.face->size() / 9 && !on; t++) on = in_element(&(*f.face)[9 * t], pr);
    if (!on) return false;
    for (int q = 0; q < 3; q++)
        if (!(x[q] >= f.bblo[q] && x[q] <= f.bbhi[q])) return false;
    return true;
}
}  // namespace

extern "C" {

// ms: per template a handle of it_create, or null (NULL: all spheres)
void *ih_create(const dem_insert_pack_spec *s, const double *lo, const double *hi, const int *periodic,
                long long ntimestep, void **ms, char *err, int errlen)
{
    auto *r = new Run();
    std::vector<dem_ins::TemplateMultisphere *> tm(s->ntemplates, nullptr);
    for (int i = 0; ms && i < s->ntemplates; i++) tm[i] = static_cast<dem_ins::TemplateMultisphere *>(ms[i]);
    auto fix = std::make_unique<dem_ins::InsertPack>();
    const std::string e = fix->init(*s, lo, hi, ntimestep, tm);
    r->fix = std::move(fix);
    if (!e.empty()) {
        std::strncpy(err, e.c_str(), errlen - 1);
        err[errlen - 1] = 0;
        delete r;
        return nullptr;
    }
    for (int d = 0; d < 3; d++) {
        r->lo[d] = lo[d];
        r->hi[d] = hi[d];
        r->periodic[d] = periodic[d];
    }
    return r;
}

// fix insert/stream with timestep dt
void *ih_stream_create(const dem_insert_stream_spec *s, const double *lo, const double *hi, const int *periodic,
                       long long ntimestep, double dt, void **ms, char *err, int errlen)
{
    auto *r = new Run();
    std::vector<dem_ins::TemplateMultisphere *> tm(s->ntemplates, nullptr);
    for (int i = 0; ms && i < s->ntemplates; i++) tm[i] = static_cast<dem_ins::TemplateMultisphere *>(ms[i]);
    auto fix = std::make_unique<dem_ins::InsertStream>();
    const std::string e = fix->init(*s, lo, hi, ntimestep, dt, tm);
    r->fix = std::move(fix);
    if (!e.empty()) {
        std::strncpy(err, e.c_str(), errlen - 1);
        err[errlen - 1] = 0;
        delete r;
        return nullptr;
    }
    for (int d = 0; d < 3; d++) {
        r->lo[d] = lo[d];
        r->hi[d] = hi[d];
        r->periodic[d] = periodic[d];
    }
    return r;
}

void ih_destroy(void *h) { delete static_cast<Run *>(h); }

long long ih_next_step(void *h) { return static_cast<Run *>(h)->fix->next_step(); }

void ih_vector(void *h, double *out) { static_cast<Run *>(h)->fix->vector(out); }

// one pre_exchange at `step`; *nbatch counts the overlap calls; rel (14 per
// atom, may be null): a stream's release data; maxrad: a radius outside the
// fix that the stream's nearby test sees
int ih_insert(void *h, long long step, int n, const double *x, const double *rad, const double *mass, int cap,
              double *xo, double *vo, double *wo, double *ro, double *mo, int *to, double *bo, int *nbatch,
              double *rel, double maxrad)
{
    Run *r = static_cast<Run *>(h);
    dem_ins::Hooks hk;
    hk.maxrad_sim = std::max(maxrad, r->fix->max_radius());
    for (int i = 0; i < n; i++) hk.maxrad_sim = std::max(hk.maxrad_sim, rad[i]);
    int calls = 0;
    hk.region_sums = [&](const double *lo, const double *hi, long long *np, double *vol, double *m) -> int {
        *np = 0;
        *vol = *m = 0.0;
        const double f4pi3 = 4. * M_PI / 3.;
        for (int i = 0; i < n; i++) {
            const double *p = x + 3 * i;
            if (p[0] >= lo[0] && p[0] <= hi[0] && p[1] >= lo[1] && p[1] <= hi[1] && p[2] >= lo[2] && p[2] <= hi[2]) {
                (*np)++;
                *vol += f4pi3 * rad[i] * rad[i] * rad[i];
                *m += mass[i];
            }
        }
        return 0;
    };
    hk.overlap = [&](const std::vector<double> &cand, const std::vector<double> &rr, const dem_ins::FaceFilter *near,
                     std::vector<uint32_t> &hit) -> int {
        calls++;
        const size_t K = cand.size() / 3;
        hit.assign(K, 0u);
        double prd[3];
        for (int d = 0; d < 3; d++) prd[d] = r->hi[d] - r->lo[d];
        for (size_t k = 0; k < K; k++) {
            const double *c = &cand[3 * k];
            uint32_t b = 0u;
            for (int j = 0; j < n; j++)
                for (int a0 = r->periodic[0] ? -1 : 0; a0 <= (r->periodic[0] ? 1 : 0); a0++)
                    for (int a1 = r->periodic[1] ? -1 : 0; a1 <= (r->periodic[1] ? 1 : 0); a1++)
                        for (int a2 = r->periodic[2] ? -1 : 0; a2 <= (r->periodic[2] ? 1 : 0); a2++) {
                            const double p0 = a0 ? x[3 * j] + a0 * prd[0] : x[3 * j];
                            const double p1 = a1 ? x[3 * j + 1] + a1 * prd[1] : x[3 * j + 1];
                            const double p2 = a2 ? x[3 * j + 2] + a2 * prd[2] : x[3 * j + 2];
                            const double pp[3] = {p0, p1, p2};
                            if (near && !nearby(*near, pp)) continue;
                            const double d0 = c[0] - p0, d1 = c[1] - p1, d2 = c[2] - p2;
                            const double rsq = d0 * d0 + d1 * d1 + d2 * d2;
                            for (size_t t = 0; t < rr.size(); t++) {
                                const double rs = rr[t] + rad[j];
                                if (rsq <= rs * rs) b |= 1u << t;
                            }
                        }
            hit[k] = b;
        }
        return 0;
    };
    std::vector<dem_ins::Placed> out;
    std::string err;
    const int rc = r->fix->insert(hk, step, out, err);
    if (nbatch) *nbatch = calls;
    if (rc) return -1;
    const int m = (int)out.size();
    for (int k = 0; k < m && k < cap; k++) {
        for (int d = 0; d < 3; d++) {
            xo[3 * k + d] = out[k].x[d];
            vo[3 * k + d] = out[k].v[d];
            wo[3 * k + d] = out[k].omega[d];
        }
        ro[k] = out[k].radius;
        mo[k] = out[k].mass;
        to[k] = out[k].type;
        if (bo) bo[k] = out[k].brid;
        if (rel && out[k].has_release) {
            const double *q = out[k].release;
            const double v[14] = {q[0], q[1], q[2], q[3], q[7], q[4], q[5], q[6], q[8], q[9], q[10], q[12], q[13], q[14]};
            std::memcpy(rel + 14 * k, v, sizeof v);
        }
    }
    return m;
}

void *it_create(const dem_template_multisphere_spec *s, char *err, int errlen)
{
    auto *t = new dem_ins::TemplateMultisphere();
    auto inside = [](const std::vector<double> &pts, const std::vector<double> &xs, const std::vector<double> &rs,
                     std::vector<uint8_t> &in) -> int {
        const size_t np = pts.size() / 3, ns = rs.size();
        in.assign(np, 0);
        for (size_t k = 0; k < np; k++)
            for (size_t j = 0; j < ns; j++) {
                double d = 0.;
                d += (pts[3 * k] - xs[3 * j]) * (pts[3 * k] - xs[3 * j]);
                d += (pts[3 * k + 1] - xs[3 * j + 1]) * (pts[3 * k + 1] - xs[3 * j + 1]);
                d += (pts[3 * k + 2] - xs[3 * j + 2]) * (pts[3 * k + 2] - xs[3 * j + 2]);
                if (d < rs[j] * rs[j]) {
                    in[k] = 1;
                    break;
                }
            }
        return 0;
    };
    const std::string e = t->init(*s, inside);
    if (!e.empty()) {
        std::strncpy(err, e.c_str(), errlen - 1);
        err[errlen - 1] = 0;
        delete t;
        return nullptr;
    }
    return t;
}

void it_destroy(void *h) { delete static_cast<dem_ins::TemplateMultisphere *>(h); }

void it_info(void *h, double *out) { static_cast<dem_ins::TemplateMultisphere *>(h)->info(out); }

void it_place(void *h, const double *x, const double *quat, double *out)
{
    auto *t = static_cast<dem_ins::TemplateMultisphere *>(h);
    double q[4] = {1.0, 0.0, 0.0, 0.0};
    if (quat) {
        const double norm = 1.0 / std::sqrt(quat[0] * quat[0] + quat[1] * quat[1] + quat[2] * quat[2] + quat[3] * quat[3]);
        for (int k = 0; k < 4; k++) q[k] = quat[k] * norm;
    }
    std::vector<double> o;
    t->place(x, q, o);
    std::memcpy(out, o.data(), o.size() * sizeof(double));
}
}
