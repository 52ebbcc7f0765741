// fix insert/pack: the reference's insertion step on the host, resident
// particles tested on the device through Hooks.  The arithmetic that decides
// what is inserted is the reference's operation for operation (built with
// -ffp-contract=off), so the inserted particles are the reference's for the
// same script and the same resident particles.
#include "insert_pack.h"

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
    s_ = s;
    if (s.seed <= 0 || s.dist_seed <= 0) return "illegal seed";
    if (s.ntemplates < 1 || s.ntemplates > DEM_MAXTYPES * 4 || !s.templates || !s.weights)
        return "illegal number of templates";
    if (s.insert_every < 0) return "must define 'insert_every'";
    if (s.maxattempt < 1) return "maxattempt must be > 0";
    if (s.ntry_mc < 1000) return "ntry_mc must be > 1000";
    if (s.vel_style < DEM_VEL_CONSTANT || s.vel_style > DEM_VEL_GAUSSIAN) return "bad vel style";
    if (s.target < DEM_INSERT_VOLUMEFRACTION || s.target > DEM_INSERT_MASS) return "bad insertion target";
    if (s.target == DEM_INSERT_VOLUMEFRACTION && !(s.target_value > 0.0 && s.target_value <= 1.0))
        return "Invalid volumefraction";
    if (s.target == DEM_INSERT_PARTICLES && !(s.target_value >= 1.0)) return "'ntotal_region' > 0 required";
    if (s.target == DEM_INSERT_MASS && !(s.target_value > 0.0)) return "'masstotal_region' > 0 required";
    for (int d = 0; d < 3; d++) {
        boxlo_[d] = boxlo[d];
        boxhi_[d] = boxhi[d];
        region_.lo[d] = s.region_lo[d];
        region_.hi[d] = s.region_hi[d];
        if (!(s.region_lo[d] < s.region_hi[d])) return "region block bounds must be lo < hi";
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
int InsertPack::x_v_omega(const Hooks &h, int ninsert, std::vector<Body> &out, std::string &err)
{
    out.clear();
    last_tries_ = 0;
    // quat_insert: identity, or the constant orientation (fix_insert.cpp:468, :316-330)
    double qins[4] = {1.0, 0.0, 0.0, 0.0};
    if (s_.orientation == DEM_ORIENT_CONSTANT)
        for (int k = 0; k < 4; k++) qins[k] = s_.quat[k];
    const bool qrand = s_.orientation == DEM_ORIENT_RANDOM;
    if (!s_.overlapcheck) {
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
        return DEM_OK;
    }
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
            do {
                gen_pos(a.x, rb);
                nt++;
            } while (s_.check_dist_from_subdomain_border && nt < maxtry && !border_ok(a.x, rb));
            a.last = nt == maxtry;
            if (!a.last) {
                gen_velocity(a.v);
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
        if (!cand.empty() && h.overlap(cand, radii_, hit)) {
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
    int ninsert_this = calc_ninsert_this(h, err);
    if (ninsert_this < 0) return DEM_ERR_ARG;
    // distribute_ninsert_this (fix_insert.cpp:893-960): a single process
    // takes all; the uncorrelated form draws one number
    if (!s_.exact_number) {
        const double fraction_local = region_volume_local_ / region_volume_;
        ninsert_this = static_cast<int>(fraction_local * static_cast<double>(ninsert_this) + random_.uniform());
    }
    ninsert_this = randomize_list(ninsert_this, step);
    last_aimed_ = ninsert_this;
    last_inserted_ = 0;
    last_tries_ = 0;
    if (ninsert_this == 0) {
        if (s_.insert_every) next_reneighbor_ += s_.insert_every;
        else next_reneighbor_ = -1;
        return DEM_OK;
    }
    // FixInsertPack::calc_maxtry (fix_insert_pack.cpp:485-490)
    if (insertion_ratio_ >= 1.) maxtry_ = ninsert_this * s_.maxattempt;
    else maxtry_ = static_cast<int>(static_cast<double>(ninsert_this * s_.maxattempt) / (1. - insertion_ratio_));
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
    last_inserted_ = (int)bodies.size();
    ninserted_ += (long long)bodies.size();
    massinserted_ += mass_this;
    if (s_.insert_every) next_reneighbor_ += s_.insert_every;
    else next_reneighbor_ = 0;
    return DEM_OK;
}

}  // namespace dem_ins
