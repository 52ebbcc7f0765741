/* pair_style gran/mi355x -- the device form of `pair_style gran`.
 *
 * Selected by `-sf mi355x` for an unchanged `pair_style gran model ...`
 * line (suffix-first lookup, force.cpp:226-238).  It parses the same model
 * keywords as PairGranProxy/the contact-model factory (pair_gran_proxy.cpp,
 * contact_models.h), registers the same history values in the same order
 * (so FixContactHistory pages and restart files keep the reference layout),
 * and collects the material tables from the reference's own property
 * registry (global_properties.cpp createYeff/createGeff/createBetaEff ...).
 * The pair computation itself runs in the device library, driven by
 * run_style verlet/mi355x; compute_force() refuses to run on the host.
 *
 * Copy into LIGGGHTS-INL src/ with the other *_mi355x files; link with
 * -lliggghts_dem (include/liggghts_dem.h).
 */
#ifdef PAIR_CLASS

PairStyle(gran/mi355x,PairGranMI355X)

#else

#ifndef LMP_PAIR_GRAN_MI355X_H
#define LMP_PAIR_GRAN_MI355X_H

#include "pair_gran.h"
#include "liggghts_dem.h"

#include <vector>

namespace LAMMPS_NS {

class PairGranMI355X : public PairGran {
 public:
  PairGranMI355X(class LAMMPS *);
  virtual ~PairGranMI355X();

  virtual void settings(int, char **);
  virtual void init_granular();
  virtual bool contact_match(const std::string mtype, const std::string model);
  virtual int64_t hashcode();
  virtual double stressStrainExponent();

  // dem_params fields of the model chain (normal, rolling, cohesion models
  // and their settings), precision of the contact law
  void fill_params(dem_params *p) const;
  // material tables of the chain into a context: dem_set_materials, and
  // dem_set_hysteretic when the normal model is hysteretic
  int upload_materials(dem_ctx *ctx) const;
  // cohesion tables (DEM_NBOND, (ntypes+1)^2 each) and settings; after the
  // state upload, as dem_set_bond requires
  int upload_bond(dem_ctx *ctx) const;
  bool has_bond_model() const { return cohesion_ != DEM_COHESION_OFF; }
  int wall_normal_model_of(const char *name) const;

 protected:
  virtual void compute_force(int eflag, int vflag);

 private:
  int normal_, rolling_, cohesion_;
  int limit_force_, tangential_damping_, torsion_torque_, mixed_;
  dem_bond_settings bond_;
  double krs_;
  // registry-connected tables ((ntypes+1)^2, index 0 unused)
  double **Yeff_, **Geff_, **betaeff_, **coeffFrict_, **coeffRollFrict_, **coeffRollVisc_;
  double **coeffRestLog_;
  double coeffRollStiffness_;

  void register_history();
  std::vector<double> flat(double **m) const;
  std::vector<double> pair_table(const char *name, bool required) const;
};

}

#endif
#endif
