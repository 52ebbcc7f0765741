// fix insert/pack: the reference's insertion step on the host, resident
// particles tested on the device through Hooks.  The arithmetic that decides
// what is inserted is the reference's operation for operation (built with
// -ffp-contract=off), so the inserted particles are the reference's for the
// same script and the same resident particles.
#include "insert_pack.h"

#include "mesh_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <unordered_map>

namespace dem_ins {

namespace {
constexpr int IA = 16807;
constexpr int IM = 2147483647;
constexpr double AM = 1.0 / IM;
constexpr int IQ = 127773;
constexpr int IR = 2836;
constexpr int kSeedOffset = 12;   // FixInsertPack's SEED_OFFSET (fix_insert_pack.cpp:67)
}  // namespace

double RanPark::uniform()
{
    const int k = seed / IQ;
    seed = IA * (seed - k * IQ) - IR * k;
    if (seed < 0) seed += IM;
    return AM * seed;
}

double RanPark::gaussian()
{
    double first;
    if (!save) {
        double v1, v2, rsq;
        do {
            v1 = 2.0 * uniform() - 1.0;
            v2 = 2.0 * uniform() - 1.0;
            rsq = v1 * v1 + v2 * v2;
        } while (!(rsq < 1.0 && rsq != 0.0));
        const double fac = std::sqrt(-2.0 * std::log(rsq) / rsq);
        second = v1 * fac;
        first = v2 * fac;
        save = 1;
    } else {
        first = second;
        save = 0;
    }
    return first;
}

bool Block::inside(const double *x) const
{
    return x[0] >= lo[0] && x[0] <= hi[0] && x[1] >= lo[1] && x[1] <= hi[1] && x[2] >= lo[2] && x[2] <= hi[2];
}

bool Block::near_face(const double *x, double cut) const
{
    if (!inside(x)) return false;
    for (int d = 0; d < 3; d++) {
        if (x[d] - lo[d] < cut) return true;
        if (hi[d] - x[d] < cut) return true;
    }
    return false;
}

// ---- fix particletemplate/multiplespheres

std::string TemplateMultisphere::init(const dem_template_multisphere_spec &s, const InsideHook &inside)
{
    s_ = s;
    const int n = s.nspheres;
    if (n < 2) return "illegal number of spheres";
    if (!s.x || !s.r) return "need to define spheres for the template";
    if (s.ntry < 1000) return "ntry is too low";
    if (!(s.scale > 0.)) return "scale factor must be >0";
    if (!(s.density > 0.)) return "density must be > 0";
    if (s.seed <= 0) return "illegal seed";
    x_.assign(s.x, s.x + 3 * (size_t)n);
    r_.assign(s.r, s.r + n);
    if (s.types) types_.assign(s.types, s.types + n);
    else types_.clear();
    const double LARGE = 1e8;
    for (int d = 0; d < 3; d++) {
        x_min_[d] = LARGE;
        x_max_[d] = -LARGE;
    }
    // the file's radii and positions times scale (the coarse-graining factor
    // of this build is 1), then the bounding box
    for (int i = 0; i < n; i++) {
        if (!(r_[i] > 0.)) return "radius must be > 0";
        r_[i] *= (s.scale * 1.0);
        for (int d = 0; d < 3; d++) x_[3 * i + d] = (s.scale * 1.0) * x_[3 * i + d];
    }
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            if (x_[3 * i + d] - r_[i] < x_min_[d]) x_min_[d] = x_[3 * i + d] - r_[i];
            if (x_[3 * i + d] + r_[i] > x_max_[d]) x_max_[d] = x_[3 * i + d] + r_[i];
        }
    random_mc_.reset(s.seed);
    random_insertion_.reset(s.seed);

    // calc_bounding_sphere (fix_template_multiplespheres.cpp:321-375): 200
    // passes over the spheres in random order, each growing a sphere to
    // contain them; the smallest is kept
    r_bound_ = LARGE;
    std::vector<char> visited(n);
    for (int shuffle = 0; shuffle < 200; shuffle++) {
        std::fill(visited.begin(), visited.end(), 0);
        int isphere = -1, nvisited = 0;
        double xb[3], rb;
        while (isphere < 0 || visited[isphere] || isphere >= n)
            isphere = static_cast<int>(random_mc_.uniform() * n);
        nvisited++;
        visited[isphere] = 1;
        for (int d = 0; d < 3; d++) xb[d] = x_[3 * isphere + d];
        rb = r_[isphere];
        while (nvisited < n) {
            while (isphere < 0 || visited[isphere] || isphere >= n)
                isphere = static_cast<int>(random_mc_.uniform() * n);
            nvisited++;
            visited[isphere] = 1;
            double dv[3];
            for (int d = 0; d < 3; d++) dv[d] = x_[3 * isphere + d] - xb[d];
            const double dist = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
            if (dist + r_[isphere] > rb) {
                const double fact = (dist + r_[isphere] - rb) / (2. * dist);
                for (int d = 0; d < 3; d++) dv[d] = fact * dv[d];
                for (int d = 0; d < 3; d++) xb[d] = xb[d] + dv[d];
                rb += std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
            }
        }
        if (rb < r_bound_) {
            r_bound_ = rb;
            for (int d = 0; d < 3; d++) x_bound_[d] = xb[d];
        }
    }
    for (int i = 0; i < n; i++) {
        double t[3];
        for (int d = 0; d < 3; d++) t[d] = x_bound_[d] - x_[3 * i + d];
        if (std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) > r_bound_)
            return "Bounding sphere calculation for template failed";
    }

    // calc_center_of_mass (fix_template_multiplespheres.cpp:410-450): ntry
    // points in the bounding box; the inside tests on the device, the
    // running mean over the inside points in their order here
    std::vector<double> pts(3 * (size_t)s.ntry);
    for (int i = 0; i < s.ntry; i++) {
        pts[3 * (size_t)i] = x_min_[0] + (x_max_[0] - x_min_[0]) * random_mc_.uniform();
        pts[3 * (size_t)i + 1] = x_min_[1] + (x_max_[1] - x_min_[1]) * random_mc_.uniform();
        pts[3 * (size_t)i + 2] = x_min_[2] + (x_max_[2] - x_min_[2]) * random_mc_.uniform();
    }
    std::vector<uint8_t> in;
    if (inside(pts, x_, r_, in)) return "device inside test failed";
    int nsuccess = 0;
    double xcm[3] = {0., 0., 0.};
    for (int i = 0; i < s.ntry; i++) {
        if (!in[i]) continue;
        for (int d = 0; d < 3; d++)
            xcm[d] = (xcm[d] * static_cast<double>(nsuccess) + pts[3 * (size_t)i + d]) / static_cast<double>(nsuccess + 1);
        nsuccess++;
    }
    volume_expect_ = static_cast<double>(nsuccess) / static_cast<double>(s.ntry) * (x_max_[0] - x_min_[0])
                     * (x_max_[1] - x_min_[1]) * (x_max_[2] - x_min_[2]);
    mass_expect_ = volume_expect_ * s.density;
    for (int d = 0; d < 3; d++) xcm_[d] = xcm[d];
    // into the centre-of-mass frame
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) x_[3 * i + d] = x_[3 * i + d] - xcm[d];
    for (int d = 0; d < 3; d++) {
        x_min_[d] = x_min_[d] - xcm[d];
        x_max_[d] = x_max_[d] - xcm[d];
        x_bound_[d] = x_bound_[d] - xcm[d];
    }
    return "";
}

void TemplateMultisphere::info(double *out) const
{
    out[0] = r_bound_;
    for (int d = 0; d < 3; d++) out[1 + d] = x_bound_[d];
    out[4] = volume_expect_;
    out[5] = mass_expect_;
    for (int d = 0; d < 3; d++) out[6 + d] = xcm_[d];
}

// MathExtraLiggghts::vec_quat_rotate (math_extra_liggghts.h:457-473): q (0, v)
// q*, by MathExtra::quatquat (math_extra.h:596-602)
static void quatquat(const double *a, const double *b, double *c)
{
    c[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    c[1] = a[0] * b[1] + b[0] * a[1] + a[2] * b[3] - a[3] * b[2];
    c[2] = a[0] * b[2] + b[0] * a[2] + a[3] * b[1] - a[1] * b[3];
    c[3] = a[0] * b[3] + b[0] * a[3] + a[1] * b[2] - a[2] * b[1];
}

void TemplateMultisphere::place(const double *x, const double *quat, std::vector<double> &out) const
{
    const int n = nspheres();
    out.resize(3 * (size_t)n);
    const double qc[4] = {quat[0], -quat[1], -quat[2], -quat[3]};
    for (int j = 0; j < n; j++) {
        double rel[3];
        for (int d = 0; d < 3; d++) rel[d] = x_[3 * j + d] - x_bound_[d];
        const double vq[4] = {0., rel[0], rel[1], rel[2]};
        double t[4], r4[4];
        quatquat(quat, vq, t);
        quatquat(t, qc, r4);
        for (int d = 0; d < 3; d++) out[3 * j + d] = r4[1 + d] + x[d];
    }
}

std::string InsertPack::init(const dem_insert_pack_spec &s, const double *boxlo, const double *boxhi,
                             long long ntimestep, const std::vector<TemplateMultisphere *> &ms)
{
    if (s.insert_every < 0) return "must define 'insert_every'";
    if (s.target < DEM_INSERT_VOLUMEFRACTION || s.target > DEM_INSERT_MASS) return "bad insertion target";
    if (s.target == DEM_INSERT_VOLUMEFRACTION && !(s.target_value > 0.0 && s.target_value <= 1.0))
        return "Invalid volumefraction";
    if (s.target == DEM_INSERT_PARTICLES && !(s.target_value >= 1.0)) return "'ntotal_region' > 0 required";
    if (s.target == DEM_INSERT_MASS && !(s.target_value > 0.0)) return "'masstotal_region' > 0 required";
    for (int d = 0; d < 3; d++) {
        region_.lo[d] = s.region_lo[d];
        region_.hi[d] = s.region_hi[d];
        if (!(s.region_lo[d] < s.region_hi[d])) return "region block bounds must be lo < hi";
    }
    std::string e = init_common(s, boxlo, boxhi, ntimestep, ms);
    if (!e.empty()) return e;

    // FixInsertPack::calc_insertion_properties (fix_insert_pack.cpp:281-290):
    // Region::reset_random(seed + SEED_OFFSET), then Region::volume_mc
    // (region.cpp:633-700) on the region's stream
    reg_random_.reset(s.seed + kSeedOffset);
    // FixParticledistributionDiscrete::max_r_bound
    double maxrb = 0.0;
    for (double rb : rbound_) maxrb = std::max(maxrb, rb);
    long long n_in = 0, n_in_local = 0;
    const Block &r = region_;
    for (int i = 0; i < s.ntry_mc; i++) {
        double pos[3];
        pos[0] = r.lo[0] + reg_random_.uniform() * (r.hi[0] - r.lo[0]);
        pos[1] = r.lo[1] + reg_random_.uniform() * (r.hi[1] - r.lo[1]);
        pos[2] = r.lo[2] + reg_random_.uniform() * (r.hi[2] - r.lo[2]);
        // Domain::is_in_domain (domain_I.h:51-63): bounds inclusive
        bool in_dom = true;
        for (int d = 0; d < 3; d++) in_dom = in_dom && pos[d] >= boxlo_[d] && pos[d] <= boxhi_[d];
        if (!in_dom) continue;
        if (r.inside(pos)) {
            n_in++;
            // Domain::is_in_subdomain pads a sub-box face on the box face by
            // 1e-8, so every point of the box is in the single sub-box
            if (!s.all_in || !r.near_face(pos, maxrb)) n_in_local++;
        }
    }
    if (n_in == 0) return "Unable to calculate region volume";
    const double vol_bbox = (r.hi[0] - r.lo[0]) * (r.hi[1] - r.lo[1]) * (r.hi[2] - r.lo[2]);
    region_volume_ = static_cast<double>(n_in) / static_cast<double>(s.ntry_mc * 1) * vol_bbox;
    if (!(region_volume_ > 0.0)) return "Fix insert: Region volume calculation with MC failed";
    double vol_local = static_cast<double>(n_in_local) / static_cast<double>(s.ntry_mc) * vol_bbox;
    const double vol_local_all = vol_local;
    if (!(vol_local_all > 0.0)) return "Unable to calculate region volume (particles too large for all_in yes?)";
    vol_local *= (region_volume_ / vol_local_all);
    region_volume_local_ = vol_local;
    // the sub-domain (the box) must hold a point of the region
    for (int d = 0; d < 3; d++)
        if (!(std::max(r.lo[d], boxlo_[d]) < std::min(r.hi[d], boxhi_[d])))
            return "Impossible to generate random points on wrong sub-domain";
    if (s.all_in)
        for (int d = 0; d < 3; d++)
            if (r.hi[d] - r.lo[d] < 2. * maxrb)
                return "Impossible to generate random points within region - region too small "
                       "(smaller than twice the particle cutoff)";
    return "";
}

// FixInsert's constructor and the distribution's set-up: keywords, the
// templates' expectancies, the seeds and the first insertion step
std::string InsertPack::init_common(const dem_insert_pack_spec &s, const double *boxlo, const double *boxhi,
                                    long long ntimestep, const std::vector<TemplateMultisphere *> &ms)
{
    s_ = s;
    if (s.seed <= 0 || s.dist_seed <= 0) return "illegal seed";
    if (s.ntemplates < 1 || s.ntemplates > DEM_MAXTYPES * 4 || !s.templates || !s.weights)
        return "illegal number of templates";
    if (s.maxattempt < 1) return "maxattempt must be > 0";
    if (s.ntry_mc < 1000) return "ntry_mc must be > 1000";
    if (s.vel_style < DEM_VEL_CONSTANT || s.vel_style > DEM_VEL_GAUSSIAN) return "bad vel style";
    for (int d = 0; d < 3; d++) {
        boxlo_[d] = boxlo[d];
        boxhi_[d] = boxhi[d];
    }
    const int T = s.ntemplates;
    if (s.orientation < DEM_ORIENT_TEMPLATE || s.orientation > DEM_ORIENT_CONSTANT) return "bad orientation";
    templ_.assign(s.templates, s.templates + T);
    distweight_.assign(s.weights, s.weights + T);
    ms_ = ms;
    ms_.resize(T, nullptr);
    volexp_.resize(T);
    massexp_.resize(T);
    rbound_.resize(T);
    for (int i = 0; i < T; i++) {
        if (distweight_[i] < 0.0) return "invalid weight";
        if (ms_[i]) {
            // FixTemplateMultiplespheres: the Monte Carlo volume and mass,
            // the bounding sphere's radius
            volexp_[i] = ms_[i]->volume();
            massexp_[i] = ms_[i]->mass();
            rbound_[i] = ms_[i]->r_bound();
            continue;
        }
        const auto &t = templ_[i];
        if (!(t.radius > 0.0) || !(t.density > 0.0)) return "template radius and density must be > 0";
        if (t.atom_type < 1) return "template atom type must be >= 1";
        // FixTemplateSphere: volume_expect = cubic_expectancy * 4 pi / 3,
        // mass_expect = expectancy(density) * volume_expect
        // (fix_template_sphere.cpp:243-245; constant pdfs: mu^3 and mu)
        volexp_[i] = t.radius * t.radius * t.radius * 4. * M_PI / 3.;
        massexp_[i] = t.density * volexp_[i];
        rbound_[i] = t.radius;
    }
    // FixParticledistributionDiscrete (fix_particledistribution_discrete.cpp:126-173)
    double wsum = 0.0;
    for (int i = 0; i < T; i++) wsum += distweight_[i];
    if (!(wsum > 0.0)) return "distribution weights sum to zero";
    for (int i = 0; i < T; i++) distweight_[i] /= wsum;
    if (!s.number_based) {
        for (int i = 0; i < T; i++) distweight_[i] = distweight_[i] / massexp_[i];
        wsum = 0.0;
        for (int i = 0; i < T; i++) wsum += distweight_[i];
        for (int i = 0; i < T; i++) distweight_[i] /= wsum;
    }
    volexpect_ = massexpect_ = 0.0;
    for (int i = 0; i < T; i++) {
        volexpect_ += volexp_[i] * distweight_[i];
        massexpect_ += massexp_[i] * distweight_[i];
    }
    // templates by insertion volume, descending, by the reference's bubble
    // sort (fix_particledistribution_discrete.cpp:197-219)
    distorder_.resize(T);
    for (int i = 0; i < T; i++) distorder_[i] = i;
    {
        bool swapped;
        int n = T;
        do {
            swapped = false;
            for (int i = 0; i < T - 1; i++)
                if (volexp_[distorder_[i]] < volexp_[distorder_[i + 1]]) {
                    std::swap(distorder_[i], distorder_[i + 1]);
                    swapped = true;
                }
            n--;
        } while (swapped && n > 0);
    }
    // the distinct sphere radii: one overlap bit each
    radii_.clear();
    radius_bit_.assign(T, {});
    for (int i = 0; i < T; i++) {
        const int ns = ms_[i] ? ms_[i]->nspheres() : 1;
        for (int j = 0; j < ns; j++) {
            const double rj = ms_[i] ? ms_[i]->radius(j) : templ_[i].radius;
            int b = -1;
            for (size_t k = 0; k < radii_.size(); k++)
                if (radii_[k] == rj) b = (int)k;
            if (b < 0) {
                b = (int)radii_.size();
                radii_.push_back(rj);
            }
            radius_bit_[i].push_back(b);
        }
    }
    if (radii_.size() > 32) return "at most 32 distinct sphere radii";

    random_.reset(s.seed);
    dist_random_.reset(s.dist_seed);
    // FixInsert: first insertion at `start`, else the next step; the next
    // re-neighbouring step is forced to it (fix_insert.cpp:369-370, :458)
    next_reneighbor_ = s.start > 0 ? s.start : ntimestep + 1;
    if (next_reneighbor_ < ntimestep + 1) return "'start' step can not be before current step";
    return "";
}

int InsertPack::max_type() const
{
    int m = 0;
    for (size_t i = 0; i < templ_.size(); i++)
        if (ms_[i])
            for (int j = 0; j < ms_[i]->nspheres(); j++) m = std::max(m, ms_[i]->type(j));
        else
            m = std::max(m, templ_[i].atom_type);
    return m;
}

// the largest and smallest sphere radius of any template
double InsertPack::max_radius() const
{
    double m = 0.0;
    for (double r : radii_) m = std::max(m, r);
    return m;
}

double InsertPack::min_radius() const
{
    double m = 1e300;
    for (double r : radii_) m = std::min(m, r);
    return m;
}

void InsertPack::vector(double *out) const
{
    out[0] = (double)ninserted_;
    out[1] = massinserted_;
    out[2] = last_inserted_;
    out[3] = last_aimed_;
    out[4] = last_tries_;
    out[5] = (double)std::max<long long>(next_reneighbor_, 0);
}

// FixInsertPack::calc_ninsert_this (fix_insert_pack.cpp:333-432)
int InsertPack::calc_ninsert_this(const Hooks &h, std::string &err)
{
    long long np_region = 0;
    double vol_region = 0.0, mass_region = 0.0;
    if (h.region_sums(region_.lo, region_.hi, &np_region, &vol_region, &mass_region)) {
        err = "region sums failed";
        return -1;
    }
    int ninsert_this = 0;
    if (s_.target == DEM_INSERT_VOLUMEFRACTION) {
        const double f = s_.target_value;
        ninsert_this = static_cast<int>((f * region_volume_ - vol_region) / volexpect_ + random_.uniform());
        insertion_ratio_ = vol_region / (f * region_volume_);
    } else if (s_.target == DEM_INSERT_PARTICLES) {
        const int ntotal = (int)s_.target_value;
        ninsert_this = ntotal - (int)np_region;
        insertion_ratio_ = static_cast<double>(np_region) / static_cast<double>(ntotal);
    } else {
        const double mt = s_.target_value;
        ninsert_this = static_cast<int>((mt - mass_region) / massexpect_ + random_.uniform());
        insertion_ratio_ = mass_region / mt;
    }
    if (ninsert_this < -200000) {
        err = "overflow in particle number calculation: inserting too many particles in one step";
        return -1;
    }
    if (ninsert_this < 0) ninsert_this = 0;
    if (insertion_ratio_ < 0.) insertion_ratio_ = 0.;
    if (insertion_ratio_ > 1.) insertion_ratio_ = 1.;
    return ninsert_this;
}

// calc_ninsert_this, then distribute_ninsert_this (fix_insert.cpp:893-960):
// a single process takes all; the uncorrelated form draws one number
int InsertPack::count_this(const Hooks &h, std::string &err)
{
    int ninsert_this = calc_ninsert_this(h, err);
    if (ninsert_this < 0) return -1;
    if (!s_.exact_number) {
        const double fraction_local = region_volume_local_ / region_volume_;
        ninsert_this = static_cast<int>(fraction_local * static_cast<double>(ninsert_this) + random_.uniform());
    }
    return ninsert_this;
}

// FixInsertPack::calc_maxtry (fix_insert_pack.cpp:485-490)
int InsertPack::calc_maxtry(int ninsert) const
{
    if (insertion_ratio_ >= 1.) return ninsert * s_.maxattempt;
    return static_cast<int>(static_cast<double>(ninsert * s_.maxattempt) / (1. - insertion_ratio_));
}

// FixInsertPack::x_v_omega's attempt (fix_insert_pack.cpp:568-582)
void InsertPack::draw_position(double *pos, double rbound, int &ntry, int maxtry)
{
    do {
        gen_pos(pos, rbound);
        ntry++;
    } while (s_.check_dist_from_subdomain_border && ntry < maxtry && !border_ok(pos, rbound));
}

// quat_insert: identity, or the constant orientation (fix_insert.cpp:468,
// :316-330)
void InsertPack::initial_quat(double *q) const
{
    q[0] = 1.0;
    q[1] = q[2] = q[3] = 0.0;
    if (s_.orientation == DEM_ORIENT_CONSTANT)
        for (int k = 0; k < 4; k++) q[k] = s_.quat[k];
}

// FixParticledistributionDiscrete::randomize_list
// (fix_particledistribution_discrete.cpp:388-470): how many of each
// template, then the pti list large to small
int InsertPack::randomize_list(int ninsert, long long step)
{
    const int T = s_.ntemplates;
    std::vector<int> parttogen(T, 0);
    if (!s_.exact_number) {
        for (int i = 0; i < T; i++)
            parttogen[i] = static_cast<int>(static_cast<double>(ninsert) * distweight_[i] + dist_random_.uniform());
    } else {
        int trunc = 0;
        std::vector<double> remainder(T);
        for (int i = 0; i < T; i++) {
            parttogen[i] = static_cast<int>(static_cast<double>(ninsert) * distweight_[i]);
            trunc += parttogen[i];
            remainder[i] = static_cast<double>(ninsert) * distweight_[i] - static_cast<double>(parttogen[i]);
        }
        const int gap = ninsert - trunc;
        for (int i = 0; i < gap; i++) {
            const double r = dist_random_.uniform() * static_cast<double>(gap);
            int j = 0;
            double rsum = remainder[0];
            while (rsum < r && j < T - 1) {
                j++;
                rsum += remainder[j];
            }
            parttogen[j]++;
        }
    }
    // each template's randomize_ptilist, in template order: a bonded
    // clump template draws one bond_random_id per body
    // (fix_template_multiplespheres.cpp:650-672)
    std::vector<std::vector<double>> ids(T);
    for (int i = 0; i < T; i++)
        if (ms_[i] && ms_[i]->bonded())
            for (int k = 0; k < parttogen[i]; k++) ids[i].push_back(ms_[i]->bond_id(step));
    // the lists wired large to small
    pti_.clear();
    for (int i = 0; i < T; i++) {
        const int c = distorder_[i];
        Pti p;
        p.templ = c;
        p.brid = -1.0;
        if (ms_[c]) {
            p.type = 0;
            p.radius = ms_[c]->r_bound();
            p.density = ms_[c]->density();
            p.mass = ms_[c]->mass();
        } else {
            const auto &t = templ_[c];
            p.type = t.atom_type;
            p.radius = t.radius;
            p.density = t.density;
            // FixTemplateSphere::randomize_ptilist (fix_template_sphere.cpp:352-354)
            p.mass = p.density * (t.radius * t.radius * t.radius * 4. * M_PI / 3.);
        }
        for (int k = 0; k < parttogen[c]; k++) {
            if (!ids[c].empty()) p.brid = ids[c][k];
            pti_.push_back(p);
        }
    }
    return (int)pti_.size();
}

// Region::generate_random / generate_random_shrinkby_cut (region.cpp:530-571)
// on the sub-domain (the box) of a single process
void InsertPack::gen_pos(double *pos, double rbound)
{
    double lo[3], hi[3], diff[3];
    for (int d = 0; d < 3; d++) {
        lo[d] = std::max(region_.lo[d], boxlo_[d]);
        hi[d] = std::min(region_.hi[d], boxhi_[d]);
        diff[d] = hi[d] - lo[d];
    }
    do {
        pos[0] = lo[0] + reg_random_.uniform() * diff[0];
        pos[1] = lo[1] + reg_random_.uniform() * diff[1];
        pos[2] = lo[2] + reg_random_.uniform() * diff[2];
    } while (!region_.inside(pos) || (s_.all_in && region_.near_face(pos, rbound)));
}

// FixInsert::generate_random_velocity (fix_insert.cpp:1025-1040)
void InsertPack::gen_velocity(double *v)
{
    for (int d = 0; d < 3; d++) v[d] = s_.vel[d];
    if (s_.vel_style == DEM_VEL_UNIFORM) {
        v[0] = s_.vel[0] + s_.vel_fluct[0] * 2.0 * (random_.uniform() - 0.50);
        v[1] = s_.vel[1] + s_.vel_fluct[1] * 2.0 * (random_.uniform() - 0.50);
        v[2] = s_.vel[2] + s_.vel_fluct[2] * 2.0 * (random_.uniform() - 0.50);
    } else if (s_.vel_style == DEM_VEL_GAUSSIAN) {
        v[0] = s_.vel[0] + s_.vel_fluct[0] * random_.gaussian();
        v[1] = s_.vel[1] + s_.vel_fluct[1] * random_.gaussian();
        v[2] = s_.vel[2] + s_.vel_fluct[2] * random_.gaussian();
    }
}

// Domain::check_dist_subbox_borders (domain_I.h:127-149), sub-box = box
bool InsertPack::border_ok(const double *pos, double rbound) const
{
    for (int d = 0; d < 3; d++) {
        const double dlo = std::fabs(boxlo_[d] - pos[d]);
        const double dhi = std::fabs(boxhi_[d] - pos[d]);
        if (dlo < rbound || dhi < rbound) return false;
    }
    return true;
}

namespace {
// The spheres placed in this insertion, binned for the overlap test among
// themselves (the reference adds each to its RegionNeighborList).
struct NewGrid {
    double h = 1.0;
    std::unordered_map<int64_t, std::vector<int>> cells;
    std::vector<double> xs, rs;
    static int64_t key(int64_t a, int64_t b, int64_t c) { return (a * 2097152 + b) * 2097152 + c; }
    int64_t ci(double v) const { return (int64_t)std::floor(v / h) + 1048576; }
    void add(const double *x, double r)
    {
        const int idx = (int)rs.size();
        xs.insert(xs.end(), x, x + 3);
        rs.push_back(r);
        cells[key(ci(x[0]), ci(x[1]), ci(x[2]))].push_back(idx);
    }
    bool overlaps(const double *x, double radius) const
    {
        const int64_t c0 = ci(x[0]), c1 = ci(x[1]), c2 = ci(x[2]);
        for (int64_t a = c0 - 1; a <= c0 + 1; a++)
            for (int64_t b = c1 - 1; b <= c1 + 1; b++)
                for (int64_t c = c2 - 1; c <= c2 + 1; c++) {
                    auto it = cells.find(key(a, b, c));
                    if (it == cells.end()) continue;
                    for (int j : it->second) {
                        double del[3];
                        del[0] = x[0] - xs[3 * j];
                        del[1] = x[1] - xs[3 * j + 1];
                        del[2] = x[2] - xs[3 * j + 2];
                        const double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
                        const double radsum = radius + rs[j];
                        if (rsq <= radsum * radsum) return true;
                    }
                }
        return false;
    }
};

// MathExtraLiggghts::random_unit_quat (math_extra_liggghts.h:616-629)
void random_unit_quat(RanPark &r, double *q)
{
    const double u1 = r.uniform();
    const double u2 = r.uniform();
    const double u3 = r.uniform();
    const double h1 = std::sqrt(1. - u1);
    const double h2 = std::sqrt(u1);
    q[0] = h1 * std::sin(2. * M_PI * u2);
    q[1] = h1 * std::cos(2. * M_PI * u2);
    q[2] = h2 * std::sin(2. * M_PI * u3);
    q[3] = h2 * std::cos(2. * M_PI * u3);
}
}  // namespace

// The body's spheres at x with orientation q: a sphere at x itself
// (particleToInsert.cpp:252), a clump's at x + rotate(x_j - x_bound, q)
// (particleToInsert.cpp:286-296)
void InsertPack::spheres_of(const Pti &p, const double *x, const double *q, std::vector<double> &xs) const
{
    if (ms_[p.templ]) {
        ms_[p.templ]->place(x, q, xs);
        return;
    }
    xs.assign(x, x + 3);
}

// FixInsertPack::x_v_omega (fix_insert_pack.cpp:498-605).  With the overlap
// check, attempts are drawn ahead in batches from copies of the streams and
// tested on the device together; the serial walk over them then takes the
// reference's decisions in order, and the streams are set to the state after
// the last attempt it used.  A batch is drawn for one template: after an
// insertion whose next body is of another template the rest of the batch is
// dropped and drawn again (positions depend on the bounding radius, the
// spheres tested on the template).
// FixInsertPack::x_v_omega without overlap check (fix_insert_pack.cpp:
// 513-545): every body placed
void InsertPack::place_unchecked(int ninsert, std::vector<Body> &out)
{
    double qins[4];
    initial_quat(qins);
    const bool qrand = s_.orientation == DEM_ORIENT_RANDOM;
    for (int k = 0; k < ninsert; k++) {
        const Pti &p = pti_[out.size()];
        Body b;
        b.pti = k;
        gen_pos(b.x, rbound_[p.templ]);
        gen_velocity(b.v);
        if (qrand) random_unit_quat(random_, qins);
        for (int d = 0; d < 4; d++) b.q[d] = qins[d];
        out.push_back(b);
    }
}

int InsertPack::x_v_omega(const Hooks &h, int ninsert, std::vector<Body> &out, std::string &err)
{
    out.clear();
    last_tries_ = 0;
    double qins[4];
    initial_quat(qins);
    const bool qrand = s_.orientation == DEM_ORIENT_RANDOM;
    if (!s_.overlapcheck) {
        place_unchecked(ninsert, out);
        return DEM_OK;
    }
    const FaceFilter *near = near_filter(h);
    NewGrid ng;
    ng.h = 2.0 * max_radius();
    int ntry = 0;
    const int maxtry = maxtry_;
    std::vector<Attempt> batch;
    std::vector<double> cand, xs;
    std::vector<uint32_t> hit;
    while (ntry < maxtry && (int)out.size() < ninsert) {
        const Pti &p0 = pti_[out.size()];
        const int templ = p0.templ;
        const double rb = rbound_[templ];
        const int ns = nsph(p0);
        // draw a batch of attempts from copies of the streams
        const int left = ninsert - (int)out.size();
        const int want = (int)std::min<long long>(std::max(1024, 4 * left), std::max(1, (1 << 20) / ns));
        batch.clear();
        cand.clear();
        RanPark reg = reg_random_, fix = random_;
        int nt = ntry;
        double q[4] = {qins[0], qins[1], qins[2], qins[3]};
        while ((int)batch.size() < want && nt < maxtry) {
            Attempt a;
            std::swap(reg_random_, reg);
            std::swap(random_, fix);
            draw_position(a.x, rb, nt, maxtry);
            a.last = nt == maxtry;
            if (!a.last) {
                attempt_velocity(a.v);
                if (qrand) random_unit_quat(random_, q);
            }
            std::swap(reg_random_, reg);
            std::swap(random_, fix);
            for (int d = 0; d < 4; d++) a.q[d] = q[d];
            a.reg = reg;
            a.fix = fix;
            a.ntry = nt;
            a.cand = (int)(cand.size() / 3);
            batch.push_back(a);
            if (a.last) break;
            spheres_of(p0, a.x, a.q, xs);
            cand.insert(cand.end(), xs.begin(), xs.end());
        }
        hit.assign(cand.size() / 3, 0u);
        if (!cand.empty() && h.overlap(cand, radii_, near, hit)) {
            err = "device overlap test failed";
            return DEM_ERR_HIP;
        }
        // the serial walk
        size_t k = 0;
        bool redraw = false;
        for (; k < batch.size(); k++) {
            const Attempt &a = batch[k];
            if (a.last) break;
            const Pti &p = pti_[out.size()];
            bool taken = false;
            for (int j = 0; j < ns && !taken; j++) taken = (hit[a.cand + j] >> radius_bit_[templ][j]) & 1u;
            for (int j = 0; j < ns && !taken; j++) taken = ng.overlaps(&cand[3 * (size_t)(a.cand + j)], sph_radius(p, j));
            if (taken) continue;
            Body b;
            b.pti = (int)out.size();
            for (int d = 0; d < 3; d++) {
                b.x[d] = a.x[d];
                b.v[d] = a.v[d];
            }
            for (int d = 0; d < 4; d++) b.q[d] = a.q[d];
            out.push_back(b);
            for (int j = 0; j < ns; j++) ng.add(&cand[3 * (size_t)(a.cand + j)], sph_radius(p, j));
            if ((int)out.size() == ninsert || pti_[out.size()].templ != templ) {
                redraw = true;
                break;
            }
        }
        const Attempt &used = batch[std::min(k, batch.size() - 1)];
        reg_random_ = used.reg;
        random_ = used.fix;
        if (qrand)
            for (int d = 0; d < 4; d++) qins[d] = used.q[d];
        ntry = used.ntry;
        if (!redraw && k < batch.size() && batch[k].last) break;   // maxtry reached
    }
    last_tries_ = ntry;
    return DEM_OK;
}

// FixInsert::pre_exchange (fix_insert.cpp:676-888) on one process
int InsertPack::insert(const Hooks &h, long long step, std::vector<Placed> &out, std::string &err)
{
    out.clear();
    if (next_reneighbor_ != step || most_recent_ == step) return DEM_OK;
    most_recent_ = step;
    int ninsert_this = count_this(h, err);
    if (ninsert_this < 0) return DEM_ERR_ARG;
    ninsert_this = randomize_list(ninsert_this, step);
    last_aimed_ = ninsert_this;
    last_inserted_ = 0;
    last_tries_ = 0;
    if (ninsert_this == 0) {
        if (s_.insert_every && more_to_insert()) next_reneighbor_ += s_.insert_every;
        else if (s_.insert_every == 0) next_reneighbor_ = -1;
        return DEM_OK;
    }
    maxtry_ = calc_maxtry(ninsert_this);
    std::vector<Body> bodies;
    const int rc = x_v_omega(h, ninsert_this, bodies, err);
    if (rc) return rc;
    // FixParticledistributionDiscrete::insert -> ParticleToInsert::insert
    // (particleToInsert.cpp:97-150): the bodies' spheres in order; a single
    // sphere has the body's mass, a clump's sphere 4.18879020479 r^3 rho
    std::vector<double> xs;
    double mass_this = 0.0;
    for (const Body &b : bodies) {
        const Pti &p = pti_[b.pti];
        mass_this += p.mass;
        spheres_of(p, b.x, b.q, xs);
        const int ns = nsph(p);
        for (int j = 0; j < ns; j++) {
            Placed a;
            for (int d = 0; d < 3; d++) {
                a.x[d] = xs[3 * j + d];
                a.v[d] = b.v[d];
                a.omega[d] = s_.omega[d];
            }
            a.radius = sph_radius(p, j);
            if (ms_[p.templ]) {
                a.mass = 4.18879020479 * a.radius * a.radius * a.radius * p.density;
                a.type = ms_[p.templ]->type(j);
            } else {
                a.mass = p.mass;
                a.type = p.type;
            }
            a.brid = p.brid;
            out.push_back(a);
        }
    }
    finalize(step, out);
    last_inserted_ = (int)bodies.size();
    ninserted_ += (long long)bodies.size();
    massinserted_ += mass_this;
    if (s_.insert_every && more_to_insert()) next_reneighbor_ += s_.insert_every;
    else next_reneighbor_ = 0;
    return DEM_OK;
}

// ---- fix insert/stream

namespace {
constexpr double kStreamTiny = 1e-14;   // FIX_INSERT_STREAM_TINY (fix_insert_stream.cpp:73)
constexpr double kSmallDmbrdr = 1.0e-8; // SMALL_DMBRDR (domain.h:64)

double mag3(const double *v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// SurfaceMesh::searchElementByAreaAcc (surface_mesh_I.h:281-295)
int search_area(const std::vector<double> &acc, double area, int lo, int hi)
{
    for (;;) {
        if ((lo < 1 || area > acc[lo - 1]) && (area <= acc[lo])) return lo;
        if ((hi < 1 || area > acc[hi - 1]) && (area <= acc[hi])) return hi;
        const int mid = static_cast<int>((lo + hi) / 2);
        if (area > acc[mid]) lo = mid;
        else hi = mid;
    }
}
}  // namespace

std::string InsertStream::init(const dem_insert_stream_spec &ss, const double *boxlo, const double *boxhi,
                               long long ntimestep, double dt, const std::vector<TemplateMultisphere *> &ms)
{
    // the keywords FixInsert parses, in the shared spec
    dem_insert_pack_spec c{};
    c.seed = ss.seed;
    c.dist_seed = ss.dist_seed;
    c.ntemplates = ss.ntemplates;
    c.templates = ss.templates;
    c.weights = ss.weights;
    c.number_based = ss.number_based;
    c.maxattempt = ss.maxattempt;
    c.insert_every = ss.insert_every;
    c.start = ss.start;
    c.overlapcheck = ss.overlapcheck;
    c.all_in = ss.all_in;
    c.check_dist_from_subdomain_border = 1;
    c.ntry_mc = ss.ntry_mc;
    c.exact_number = ss.exact_number;
    c.vel_style = ss.vel_style;
    for (int d = 0; d < 3; d++) {
        c.vel[d] = ss.vel[d];
        c.vel_fluct[d] = ss.vel_fluct[d];
        c.omega[d] = ss.omega[d];
    }
    c.group_mask = ss.group_mask;
    c.multisphere = ss.multisphere;
    c.orientation = ss.orientation;
    for (int k = 0; k < 4; k++) c.quat[k] = ss.quat[k];
    if (ss.orientation < DEM_ORIENT_TEMPLATE || ss.orientation > DEM_ORIENT_CONSTANT) return "bad orientation";
    if (ss.insert_every < -1) return "'insert_every' must be > 0";
    if (ss.duration < 0) return "'duration' can not be < 1";
    if (!(dt > 0.0)) return "set the timestep first";
    std::string e = init_common(c, boxlo, boxhi, ntimestep, ms);
    if (!e.empty()) return e;
    dt_ = dt;
    parallel_ = ss.parallel != 0;
    duration_ = ss.duration;
    extrude_ = ss.extrude_length;
    if (extrude_ < 0.) return "invalid extrude_length";
    // nparticles / mass (fix_insert.cpp:132-145): INF leaves the total open
    ninsert_exists_ = !(ss.nparticles < 0 || ss.mass < 0.0);
    ninsert_ = ss.nparticles > 0 ? ss.nparticles : 0;
    massinsert_ = ss.mass > 0.0 ? ss.mass : 0.0;
    nflowrate_ = ss.particlerate;
    massflowrate_ = ss.massrate;
    // FixInsert::sanity_check (fix_insert.cpp:485-512)
    if (ninsert_ > 0 && massinsert_ > 0.) return "must not define both 'nparticles' and 'mass'";
    if (nflowrate_ > 0. && massflowrate_ > 0.) return "must not define both 'massrate' and 'particlerate'";
    if (ss.insert_every == 0 && (massflowrate_ > 0. || nflowrate_ > 0.))
        return "must not define 'particlerate' or 'massrate' for 'insert_every' = 0";

    // the insertion face: fix mesh/surface's element properties
    if (ss.face_ntri < 1 || !ss.face_nodes) return "must define an insertion face";
    if (ss.all_in) return "using all_in yes requires you to use a fix mesh/surface/planar";
    dem::MeshHost m;
    dem::mesh_build(m, ss.face_ntri, ss.face_nodes, 0.0, boxlo, boxhi);
    if (m.ntri < 1) return "must define an insertion face";
    // SurfaceMesh::isPlanar (surface_mesh_I.h:1291-1313): every pair of
    // neighbours coplanar, |n_a . n_b| > curvature (areCoplanarNeighs, :894-917)
    for (int t = 0; t < m.ntri; t++)
        for (int k = 0; k < m.nneigh[t]; k++) {
            const int j = m.neigh[dem::kMeshNeighMax * t + k];
            if (!(std::fabs(dot3(&m.snorm[3 * t], &m.snorm[3 * j])) > m.curvature))
                return "command requires a planar face for insertion";
        }
    face_ = m.node;
    // TriMesh::calcArea (tri_mesh_I.h:327-338) and the accumulated areas of
    // recalcLocalSurfProperties (surface_mesh_I.h:187-217)
    area_acc_.assign(m.ntri, 0.0);
    area_ = 0.0;
    for (int t = 0; t < m.ntri; t++) {
        const double *e0 = &m.evec[9 * t], *e1 = e0 + 3;
        double cr[3];
        cr[0] = e0[1] * e1[2] - e0[2] * e1[1];
        cr[1] = e0[2] * e1[0] - e0[0] * e1[2];
        cr[2] = e0[0] * e1[1] - e0[1] * e1[0];
        const double a = 0.5 * mag3(cr) * m.elen[3 * t] * m.elen[3 * t + 1];
        area_acc_[t] = a;
        if (t > 0) area_acc_[t] += area_acc_[t - 1];
        area_ += a;
    }
    // the mesh's stream (MultiNodeMesh, multi_node_mesh_I.h:61)
    reg_random_.reset(179424799);

    // FixInsertStream::calc_insertion_properties (fix_insert_stream.cpp:287-423)
    for (int d = 0; d < 3; d++) nrm_[d] = m.snorm[d];
    double dot = dot3(s_.vel, nrm_);
    if (dot < 0)
        for (int d = 0; d < 3; d++) nrm_[d] = nrm_[d] * -1.;
    dot = dot3(s_.vel, nrm_);
    for (int d = 0; d < 3; d++) vn_[d] = nrm_[d] * dot;
    const double vnmag = mag3(vn_);
    if (vnmag < 1.e-3) return "insertion velocity projected on face normal is < 1e-3";
    for (int d = 0; d < 3; d++) pref_[d] = face_[d];
    if (mag3(s_.vel) < 1e-5) return "insertion velocity too low";
    double maxrb = 0.0;
    for (double rb : rbound_) maxrb = std::max(maxrb, rb);
    if (s_.insert_every == -1 && extrude_ == 0.) return "must define either 'insert_every' or 'extrude_length'";
    if (s_.insert_every > -1 && extrude_ > 0.) return "must not provide both 'insert_every' and 'extrude_length'";
    if (extrude_ > 0. && duration_ > 0) return "must not provide both 'extrude_length' and 'duration'";
    if (s_.insert_every == -1) {
        if (extrude_ < 3. * maxrb && (s_.all_in || s_.overlapcheck)) return "'extrude_length' is too small";
        s_.insert_every = static_cast<int>((extrude_ + kStreamTiny) / (dt_ * vnmag));
        if (s_.insert_every == 0) return "insertion velocity too high or extrude_length too low";
    } else {
        if (s_.insert_every < 1) return "'insert_every' must be > 0";
        if (duration_ == 0) duration_ = s_.insert_every;
        else if (duration_ > s_.insert_every) return "'duration' > 'insert_every' not allowed";
        extrude_ = static_cast<double>(duration_) * dt_ * vnmag;
        if (extrude_ < 2. * maxrb)
            return "'insert_every' or 'vel' is too small, or (bounding) radius of inserted particles too large";
    }
    if (ninsert_ == 0 && ninsert_exists_) {
        if (massinsert_ / massexpect_ > 2.e9) return "you are attempting to insert more than 2e9 particles";
        if (massinsert_ > 0.) ninsert_ = static_cast<int>((massinsert_ + kStreamTiny) / massexpect_);
        else return "must define either 'nparticles' or 'mass'";
    }
    if (nflowrate_ == 0.) {
        if (massflowrate_ == 0.) return "must define either 'massrate' or 'particlerate'";
        nflowrate_ = massflowrate_ / massexpect_;
    } else {
        massflowrate_ = nflowrate_ * massexpect_;
    }
    ninsert_per_ = nflowrate_ * (static_cast<double>(s_.insert_every) * dt_);
    if (ninsert_exists_) massinsert_ = static_cast<double>(ninsert_) * massexpect_;
    ext_min_ = 0.;
    ext_max_ = extrude_;
    return "";
}

// TriMesh::generateRandomOwnedGhost (tri_mesh_I.h:395-428): a triangle by
// accumulated area, then barycentric coordinates, on the mesh's stream
void InsertStream::face_point(double *pos)
{
    const double r = reg_random_.uniform() * (area_ + 0.0);
    const int t = search_area(area_acc_, r, 0, (int)area_acc_.size() - 1);
    double u, v;
    do {
        u = reg_random_.uniform();
        v = reg_random_.uniform();
    } while (u + v > 1);
    const double b0 = 1. - u - v, b1 = v, b2 = u;
    const double *nd = &face_[9 * (size_t)t];
    for (int d = 0; d < 3; d++) pos[d] = b0 * nd[d] + b1 * nd[3 + d] + b2 * nd[6 + d];
}

// FixInsertStream::generate_random (fix_insert_stream.cpp:617-641)
void InsertStream::gen_random(double *pos, double rad)
{
    face_point(pos);
    double r;
    if (s_.overlapcheck) r = -1. * (random_.uniform() * (ext_max_) + rad + ext_min_);
    else r = -1. * (random_.uniform() * (ext_max_ - 2. * rad) + rad + ext_min_);
    for (int d = 0; d < 3; d++) pos[d] = pos[d] + nrm_[d] * r;
}

// Domain::is_in_subdomain (domain_I.h:65-88) for the single sub-box, whose
// faces all lie on the box's and get the 1e-8 padding
bool InsertStream::in_subdomain(const double *pos) const
{
    for (int d = 0; d < 3; d++)
        if (!(pos[d] >= boxlo_[d] - kSmallDmbrdr && pos[d] < boxhi_[d] + kSmallDmbrdr)) return false;
    return true;
}

// FixInsertStream::calc_ins_fraction (fix_insert_stream.cpp:493-556), run by
// the first insertion
std::string InsertStream::calc_ins_fraction()
{
    do_fraction_ = false;
    int n_in_local = 0;
    const int n_test = s_.ntry_mc;
    for (int i = 0; i < n_test; i++) {
        double pos[3];
        face_point(pos);
        const double r = -1. * (random_.uniform() * extrude_);
        for (int d = 0; d < 3; d++) pos[d] = pos[d] + nrm_[d] * r;
        if (in_subdomain(pos)) n_in_local++;
    }
    ins_fraction_ = static_cast<double>(n_in_local) / static_cast<double>(n_test);
    if (parallel_) {
        ext_min_ = extrude_;
        ext_max_ = 0.;
        for (int ix = 0; ix < 2; ix++)
            for (int iy = 0; iy < 2; iy++)
                for (int iz = 0; iz < 2; iz++) {
                    double be[3];
                    be[0] = (ix == 0 ? boxlo_[0] : boxhi_[0]) - pref_[0];
                    be[1] = (iy == 0 ? boxlo_[1] : boxhi_[1]) - pref_[1];
                    be[2] = (iz == 0 ? boxlo_[2] : boxhi_[2]) - pref_[2];
                    const double dot = -dot3(be, nrm_);
                    if (dot > 0. && dot < extrude_) {
                        ext_max_ = std::max(ext_max_, dot);
                        ext_min_ = std::min(ext_min_, dot);
                    } else if (dot < 0.) {
                        ext_min_ = 0.;
                    } else if (dot >= extrude_) {
                        ext_max_ = extrude_;
                    }
                }
        if (ext_min_ == extrude_) ext_min_ = 0.;
        if (ext_max_ == 0.) ext_max_ = extrude_;
    }
    if (ins_fraction_ < 0.9 || ins_fraction_ > 1.1)
        return "insertion volume could not be distributed properly in parallel. Bad decomposition or insertion "
               "face extrusion is too small or outside domain";
    return "";
}

// FixInsert::calc_ninsert_this (fix_insert.cpp:658-668), pre_exchange's cap
// (:693-700) and distribute_ninsert_this (:893-960) on one process
int InsertStream::count_this(const Hooks &, std::string &err)
{
    if (ninsert_per_ == 0.) {
        err = "ninsert_per == 0.";
        return -1;
    }
    int n = static_cast<int>(ninsert_per_ + random_.uniform());
    if (ninsert_exists_ && ninserted_ + n > ninsert_) n = (int)(ninsert_ - ninserted_);
    if (ninsert_exists_ && ninserted_ + n >= ninsert_) {
        n = (int)(ninsert_ - ninserted_);
        if (n < 0) n = 0;
        s_.exact_number = 1;
    }
    // FixInsertStream::insertion_fraction (fix_insert_stream.cpp:475-486)
    if (do_fraction_) {
        err = calc_ins_fraction();
        if (!err.empty()) return -1;
    }
    const double fraction_local = ins_fraction_;
    if (!s_.exact_number) return static_cast<int>(fraction_local * static_cast<double>(n) + random_.uniform());
    // exact: the one process's share, normalised, is all of it
    if (fraction_local < 0.02) {
        err = "Internal error distributing particles to processes";
        return -1;
    }
    return n;
}

// FixInsertStream::x_v_omega's attempt with overlap check
// (fix_insert_stream.cpp:730-739)
void InsertStream::draw_position(double *pos, double rbound, int &ntry, int maxtry)
{
    do {
        gen_random(pos, rbound);
        ntry++;
    } while (ntry < maxtry && (!in_subdomain(pos) || !border_ok(pos, rbound)));
}

// set_x_v_omega gets v_normal; the insertion velocity waits for the release
void InsertStream::attempt_velocity(double *v)
{
    for (int d = 0; d < 3; d++) v[d] = vn_[d];
}

// FixInsertStream::x_v_omega without overlap check (fix_insert_stream.cpp:
// 684-716): once maxtry is reached, every remaining body still draws one
// position
void InsertStream::place_unchecked(int ninsert, std::vector<Body> &out)
{
    double qins[4];
    initial_quat(qins);
    const bool qrand = s_.orientation == DEM_ORIENT_RANDOM;
    int ntry = 0;
    const int maxtry = maxtry_;
    for (int it = 0; it < ninsert; it++) {
        const Pti &p = pti_[out.size()];
        double pos[3];
        do {
            gen_random(pos, rbound_[p.templ]);
            ntry++;
        } while (ntry < maxtry && !in_subdomain(pos));
        if (ntry < maxtry) {
            if (qrand) random_unit_quat(random_, qins);
            Body b;
            b.pti = (int)out.size();
            for (int d = 0; d < 3; d++) {
                b.x[d] = pos[d];
                b.v[d] = vn_[d];
            }
            for (int d = 0; d < 4; d++) b.q[d] = qins[d];
            out.push_back(b);
        }
    }
    last_tries_ = ntry;
}

// FixInsertStream::is_nearby and getBoundingBox (fix_insert_stream.cpp:
// 568-610) for this insertion
const FaceFilter *InsertStream::near_filter(const Hooks &h)
{
    const double maxrad = h.maxrad_sim;
    for (int d = 0; d < 3; d++) {
        near_.pref[d] = pref_[d];
        near_.nrm[d] = nrm_[d];
    }
    near_.dmax = maxrad;
    near_.dmin = -(extrude_ + 3. * maxrad);
    double lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
        lo[d] = face_[d];
        hi[d] = face_[d];
    }
    for (size_t k = 0; k < face_.size(); k++) {
        lo[k % 3] = std::min(lo[k % 3], face_[k]);
        hi[k % 3] = std::max(hi[k % 3], face_[k]);
    }
    const double cut = 3 * maxrad;
    const double delta = -(extrude_ + 2 * cut);
    double maxrb = 0.0;
    for (double rb : rbound_) maxrb = std::max(maxrb, rb);
    const double extend = 3 * maxrad + 2. * maxrb;
    for (int d = 0; d < 3; d++) {
        // BoundingBox::extrude, shrinkToSubbox, extendByDelta
        // (bounding_box.cpp:77-95, bounding_box.h:139-155)
        lo[d] = std::min(lo[d], lo[d] + delta * nrm_[d]);
        hi[d] = std::max(hi[d], hi[d] + delta * nrm_[d]);
        if (lo[d] < boxlo_[d]) lo[d] = boxlo_[d];
        if (hi[d] > boxhi_[d]) hi[d] = boxhi_[d];
        near_.bblo[d] = lo[d] - extend;
        near_.bbhi[d] = hi[d] + extend;
    }
    near_.face = &face_;
    return &near_;
}

// FixInsertStream::finalize_insertion (fix_insert_stream.cpp:763-826): the
// release data of every new atom, its insertion velocity drawn in atom order
void InsertStream::finalize(long long step, std::vector<Placed> &out)
{
    const double vnmag = mag3(vn_);
    const int istep = static_cast<int>(step);
    for (Placed &a : out) {
        double rel[3];
        for (int d = 0; d < 3; d++) rel[d] = pref_[d] - a.x[d];
        const double dist_normal = dot3(rel, nrm_);
        const int n_steps = static_cast<int>((dist_normal + kStreamTiny) / (vnmag * dt_));
        double vins[3];
        gen_velocity(vins);
        double *r = a.release;
        for (int d = 0; d < 3; d++) {
            r[d] = a.x[d];
            r[4 + d] = vn_[d];
            r[8 + d] = vins[d];
            r[12 + d] = s_.omega[d];
            a.omega[d] = 0.0;
        }
        r[3] = static_cast<double>(istep);
        r[7] = static_cast<double>(istep + n_steps);
        r[11] = 0.0;
        r[15] = 0.0;
        a.has_release = true;
    }
}

}  // namespace dem_ins
