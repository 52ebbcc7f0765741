// fix insert/pack on the device's particles: the reference's placement
// loop (FixInsertPack::x_v_omega, fix_insert_pack.cpp:498-605) and its
// random streams on the host, the overlap test of each batch of candidate
// positions against the resident particles on the device (hooks supplied by
// the engine).
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/liggghts_dem.h"

namespace dem_ins {

// RanPark (random_park.cpp:44-103): Park and Miller's minimal standard
// generator, seed <- 16807 seed mod (2^31 - 1) by Schrage's factorisation;
// gaussian() is the polar Box-Muller method keeping the second value.
struct RanPark {
    int seed = 1;
    int save = 0;
    double second = 0.0;
    double uniform();
    double gaussian();
    void reset(int s)
    {
        seed = s;
        save = 0;
    }
};

// A block region's bounds (region_block.cpp), units box.
struct Block {
    double lo[3], hi[3];
    // RegBlock::inside (region_block.cpp:272-277): bounds inclusive
    bool inside(const double *x) const;
    // RegBlock::surface_interior (region_block.cpp:286-345) > 0: inside and
    // closer than cut to a face
    bool near_face(const double *x, double cut) const;
};

// One body of the pti list (ParticleToInsert): a sphere of a
// particletemplate/sphere, or a clump of a particletemplate/multiplespheres.
struct Pti {
    int templ;        // index in the spec's template order
    int type;
    double radius;    // a sphere's radius; a clump's bounding radius
    double density;
    double mass;      // mass_ins: density * r^3 * 4 pi / 3 (fix_template_sphere.cpp:352-354),
                      // or the clump's Monte Carlo mass
    double brid;      // bond_random_id of a bonded clump, -1 else
};

// One created atom.
struct Placed {
    double x[3], v[3], omega[3];
    double radius, mass;
    int type;
    double brid;
    // fix insert/stream's release data: (x_ins, i_step), (v_integrate,
    // r_step), (v_release, 0), (omega_release, 0)
    bool has_release = false;
    double release[16];
};

// The resident particles a fix insert/stream tests against: normal distance
// d = (x - pref).nrm within [dmin, dmax], x + d nrm on a face triangle (9
// doubles each in face), x inside [bblo, bbhi] (FixInsertStream::is_nearby,
// fix_insert_stream.cpp:568-593; FixInsert::load_xnear, fix_insert.cpp:
// 980-1018).
struct FaceFilter {
    double pref[3], nrm[3];
    double dmin, dmax;
    double bblo[3], bbhi[3];
    const std::vector<double> *face;
};

// What the placement needs from the resident particles.
struct Hooks {
    // Modify::get_max_radius (modify_liggghts.cpp:575-620): the largest
    // radius of any atom or insertion template
    double maxrad_sim = 0.0;
    // particles with position inside [lo, hi] (inclusive): count, sum of
    // 4/3 pi r^3, sum of mass (FixInsertPack::calc_ninsert_this)
    std::function<int(const double *lo, const double *hi, long long *np, double *vol, double *mass)> region_sums;
    // bit t of hit[k]: a sphere of radius rad[t] at cand[3k..3k+2] overlaps
    // a resident particle, rsq <= (rad[t] + r_j)^2 (RegionNeighborList::
    // hasOverlap, region_neighbor_list_I.h:72-98), periodic images
    // included; near != null: only the residents it keeps
    std::function<int(const std::vector<double> &cand, const std::vector<double> &rad, const FaceFilter *near,
                      std::vector<uint32_t> &hit)>
        overlap;
};

// fix particletemplate/multiplespheres (fix_template_multiplespheres.cpp):
// a clump of spheres with constant density, its bounding sphere (randomised
// iterative construction, 200 shuffles) and its centre of mass (Monte Carlo
// over the bounding box), both on the template's `random_mc` stream; the
// spheres are kept relative to the centre of mass.
class TemplateMultisphere {
  public:
    // inside[k] = 1: point k lies inside a sphere, dist^2 < r^2 with the
    // squared distance summed x, y, z (FixTemplateMultiplespheres::dist_sqr)
    using InsideHook = std::function<int(const std::vector<double> &pts, const std::vector<double> &xs,
                                         const std::vector<double> &rs, std::vector<uint8_t> &inside)>;
    std::string init(const dem_template_multisphere_spec &s, const InsideHook &inside);
    // ParticleToInsert::set_x_v_omega (particleToInsert.cpp:330-350): sphere
    // j at x + rotate(x_j - x_bound, quat)
    void place(const double *x, const double *quat, std::vector<double> &out) const;
    int nspheres() const { return (int)r_.size(); }
    double radius(int j) const { return r_[j]; }
    int type(int j) const { return types_.empty() ? s_.atom_type : types_[j]; }
    double density() const { return s_.density; }
    bool bonded() const { return s_.bonded != 0; }
    // bond_random_id of a clump created at `step`: step + u (random_insertion)
    double bond_id(long long step) { return static_cast<double>(step) + random_insertion_.uniform(); }
    void info(double *out8) const;
    double r_bound() const { return r_bound_; }
    double volume() const { return volume_expect_; }
    double mass() const { return mass_expect_; }

  private:
    dem_template_multisphere_spec s_{};
    std::vector<double> x_, r_;
    std::vector<int> types_;
    double x_min_[3], x_max_[3], x_bound_[3], xcm_[3];
    double r_bound_ = 0.0, volume_expect_ = 0.0, mass_expect_ = 0.0;
    RanPark random_mc_, random_insertion_;
};

// fix insert/pack, and the FixInsert machinery fix insert/stream shares
// (the distribution, the batched placement walk, pre_exchange's bookkeeping).
class InsertPack {
  public:
    virtual ~InsertPack() = default;
    // Checks the spec, builds the distribution and runs the region volume
    // estimate; returns an error text, empty on success.
    // ms[i]: the multisphere template of template i, or null for a sphere
    std::string init(const dem_insert_pack_spec &s, const double *boxlo, const double *boxhi, long long ntimestep,
                     const std::vector<TemplateMultisphere *> &ms);
    long long next_step() const { return next_reneighbor_; }
    bool due(long long step) const { return next_reneighbor_ > 0 && step == next_reneighbor_; }
    // One FixInsert::pre_exchange: the particles to create, in creation
    // order.  Returns DEM_OK or an error code with err set.
    int insert(const Hooks &h, long long step, std::vector<Placed> &out, std::string &err);
    void vector(double *out6) const;
    int max_type() const;
    double max_radius() const;
    double min_radius() const;
    virtual bool is_stream() const { return false; }

  protected:
    struct Body {
        int pti;
        double x[3], v[3], q[4];
    };
    // the spec's template, distribution, seed and keyword fields
    std::string init_common(const dem_insert_pack_spec &s, const double *boxlo, const double *boxhi,
                            long long ntimestep, const std::vector<TemplateMultisphere *> &ms);
    // bodies to insert at this step, after distribute_ninsert_this; < 0:
    // error (err set)
    virtual int count_this(const Hooks &h, std::string &err);
    virtual int calc_maxtry(int ninsert) const;
    // FixInsert's `!ninsert_exists || ninserted < ninsert`
    virtual bool more_to_insert() const { return true; }
    // one attempt's position: draws until it passes the fix's tests or
    // ntry reaches maxtry
    virtual void draw_position(double *pos, double rbound, int &ntry, int maxtry);
    // an attempt's velocity
    virtual void attempt_velocity(double *v) { gen_velocity(v); }
    // the placement loop without overlap check
    virtual void place_unchecked(int ninsert, std::vector<Body> &out);
    virtual const FaceFilter *near_filter(const Hooks &) { return nullptr; }
    // after the atoms are created (FixInsert*::finalize_insertion)
    virtual void finalize(long long step, std::vector<Placed> &out)
    {
        (void)step;
        (void)out;
    }
    void initial_quat(double *q) const;

    struct Attempt {
        double x[3], v[3], q[4];
        RanPark reg, fix;    // the streams after this attempt
        int ntry;            // tries counted after it
        bool last;           // reached maxtry: no velocity, no test
        int cand;            // its first sphere's candidate index
    };
    int calc_ninsert_this(const Hooks &h, std::string &err);
    int randomize_list(int ntotal, long long step);
    void gen_pos(double *pos, double rbound);
    void gen_velocity(double *v);
    bool border_ok(const double *pos, double rbound) const;
    int x_v_omega(const Hooks &h, int ninsert, std::vector<Body> &out, std::string &err);
    void spheres_of(const Pti &p, const double *x, const double *q, std::vector<double> &xs) const;
    int nsph(const Pti &p) const { return ms_[p.templ] ? ms_[p.templ]->nspheres() : 1; }
    double sph_radius(const Pti &p, int j) const { return ms_[p.templ] ? ms_[p.templ]->radius(j) : p.radius; }

    dem_insert_pack_spec s_{};
    std::vector<dem_template_sphere> templ_;
    std::vector<double> distweight_, volexp_, massexp_;
    std::vector<int> distorder_;
    std::vector<TemplateMultisphere *> ms_;
    std::vector<double> rbound_;       // per template: bounding radius
    std::vector<double> radii_;        // distinct sphere radii, the overlap bit order
    std::vector<std::vector<int>> radius_bit_;   // per template and sphere: its bit
    double volexpect_ = 0.0, massexpect_ = 0.0;
    double boxlo_[3] = {0, 0, 0}, boxhi_[3] = {0, 0, 0};
    Block region_{};
    RanPark random_;                   // the fix's stream (seed)
    RanPark reg_random_;               // the region's stream (seed + 12)
    RanPark dist_random_;              // the distribution's stream
    double region_volume_ = 0.0, region_volume_local_ = 0.0;
    double insertion_ratio_ = 0.0;
    int maxtry_ = 0;
    long long next_reneighbor_ = 0;
    long long most_recent_ = -1;
    std::vector<Pti> pti_;
    // totals
    long long ninserted_ = 0;
    double massinserted_ = 0.0;
    int last_inserted_ = 0, last_aimed_ = 0, last_tries_ = 0;
};

// fix insert/stream (fix_insert_stream.cpp): bodies placed in the insertion
// face extruded against the inflow, the face's points drawn on the mesh's
// stream (held in reg_random_), released after the time they need to cross
// the face at the normal velocity.
class InsertStream : public InsertPack {
  public:
    std::string init(const dem_insert_stream_spec &s, const double *boxlo, const double *boxhi, long long ntimestep,
                     double dt, const std::vector<TemplateMultisphere *> &ms);
    bool is_stream() const override { return true; }
    int insert_every() const { return s_.insert_every; }
    double ninsert_per() const { return ninsert_per_; }
    const double *normal() const { return nrm_; }
    const double *v_normal() const { return vn_; }
    double extrude_length() const { return extrude_; }
    // the face triangles, 9 doubles each
    const std::vector<double> &face() const { return face_; }

  protected:
    int count_this(const Hooks &h, std::string &err) override;
    int calc_maxtry(int ninsert) const override { return ninsert * s_.maxattempt; }
    bool more_to_insert() const override { return !ninsert_exists_ || ninserted_ < ninsert_; }
    void draw_position(double *pos, double rbound, int &ntry, int maxtry) override;
    void attempt_velocity(double *v) override;
    void place_unchecked(int ninsert, std::vector<Body> &out) override;
    const FaceFilter *near_filter(const Hooks &h) override;
    void finalize(long long step, std::vector<Placed> &out) override;

  private:
    void face_point(double *pos);
    void gen_random(double *pos, double rad);
    bool in_subdomain(const double *pos) const;
    std::string calc_ins_fraction();
    std::vector<double> face_, area_acc_;
    double area_ = 0.0;
    double nrm_[3] = {0, 0, 0}, vn_[3] = {0, 0, 0}, pref_[3] = {0, 0, 0};
    double extrude_ = 0.0, ext_min_ = 0.0, ext_max_ = 0.0;
    double dt_ = 0.0, ninsert_per_ = 0.0, nflowrate_ = 0.0, massflowrate_ = 0.0, massinsert_ = 0.0;
    int ninsert_ = 0, duration_ = 0;
    bool ninsert_exists_ = true, parallel_ = false;
    bool do_fraction_ = true;
    double ins_fraction_ = 0.0;
    FaceFilter near_{};
};

}  // namespace dem_ins
