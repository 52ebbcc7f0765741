/* pair_style gran/mi355x: model parsing, history registration and material
 * tables for the device library.  The arithmetic of the contact models runs
 * on the GPU (liggghts-inl_amd/csrc); nothing here computes a force. */
#include "pair_gran_mi355x.h"

#include "atom.h"
#include "error.h"
#include "fix_property_global.h"
#include "force.h"
#include "global_properties.h"
#include "modify.h"
#include "property_registry.h"

#include <cstdlib>
#include <cstring>
#include <string>

using namespace LAMMPS_NS;

namespace {

// history values in the order the reference models register them
// (normal model, cohesion model, tangential model, rolling model:
// contact_models.h member order) with their newton flags
struct HistVal { const char *name, *newton; };

const HistVal kHysteretic[] = {
    {"deltaMax", "0"}, {"deltaZero", "0"}, {"k1", "0"}, {"deltaZero_old", "0"}, {"k1_old", "0"},
    {"delta_old", "0"}, {"deltaMin", "0"}, {"betan", "0"}, {"f0", "0"}, {"f0_old", "0"},
    {"kc", "1"}, {"fo", "1"}};  // normal_model_hysteretic_nonlinear1.h:82-93
const HistVal kBond[] = {
    {"contflag", "0"}, {"initial_dist", "0"}, {"contactPosX", "0"}, {"contactPosY", "0"},
    {"contactPosZ", "0"}, {"ftx", "1"}, {"fty", "1"}, {"ftz", "1"}, {"torquenx", "1"},
    {"torqueny", "1"}, {"torquenz", "1"}, {"torquetx", "1"}, {"torquety", "1"},
    {"torquetz", "1"}};  // cohesion_model_bond.h:237-250
const HistVal kBondNonlinear[] = {
    {"bondFlag", "0"}, {"initial_dist", "0"}, {"contactPosX", "0"}, {"contactPosY", "0"},
    {"contactPosZ", "0"}, {"ftx", "1"}, {"fty", "1"}, {"ftz", "1"}, {"thetanx", "1"},
    {"thetany", "1"}, {"thetanz", "1"}, {"thetatx", "1"}, {"thetaty", "1"}, {"thetatz", "1"},
    {"deltan_max", "0"}, {"thetanx_max", "0"}, {"thetany_max", "0"}, {"thetanz_max", "0"},
    {"thetatx_max", "0"}, {"thetaty_max", "0"}, {"thetatz_max", "0"}, {"thetanx_min", "0"},
    {"thetany_min", "0"}, {"thetanz_min", "0"}, {"thetatx_min", "0"}, {"thetaty_min", "0"},
    {"thetatz_min", "0"}, {"deltan_min", "0"}};  // cohesion_model_bond_nonlinear.h:204-231
const HistVal kShear[] = {{"shearx", "1"}, {"sheary", "1"}, {"shearz", "1"}};  // tangential_model_history.h:79-81
const HistVal kRoll[] = {{"r_torquex_old", "1"}, {"r_torquey_old", "1"}, {"r_torquez_old", "1"}};  // rolling_model_epsd.h:71-73

// the hysteretic tables in the order dem_set_hysteretic takes them
const char *kHystProps[DEM_NHYST] = {
    "LoadingStiffness", "UnloadingStiffness", "coefficientAdhesionStiffness", "coefficientPlasticityDepth",
    "pullOffForce", "alphaCustom", "cinCustom", "aoneCustom", "atwoCustom", "athreeCustom", "kcinCustom"};

// bond tables in DEM_NBOND order: the cohesion bond property names, and the
// bond/nonlinear ones where that model names them differently
const char *kBondProps[DEM_NBOND] = {
    "radiusMultiplierBond", "normalBondStiffnessPerUnitArea", "tangentialBondStiffnessPerUnitArea",
    "dampingNormalForceBond", "dampingTangentialForceBond", "dampingNormalTorqueBond",
    "dampingTangentialTorqueBond", "maxDistanceBond", "maxSigmaBond", "maxTauBond",
    "ratioTensionCompressionBond", "createDistanceBond", "stiffnessPerUnitAreaK_fn1",
    "stiffnessPerUnitAreaKu_fn1", "stiffnessPerUnitAreaKc_fn1", "stiffnessPerUnitAreaK_fn2",
    "stiffnessPerUnitAreaKu_fn2", "stiffnessPerUnitAreaKc_fn2", "stiffnessPerUnitAreaK_ft",
    "stiffnessPerUnitAreaK_tn", "stiffnessPerUnitAreaKu_tn", "stiffnessPerUnitAreaKc_tn",
    "stiffnessPerUnitAreaK_tt", "stiffnessPerUnitAreaKu_tt", "stiffnessPerUnitAreaKc_tt"};
const char *kBondNonlinearProps[12] = {
    "radiusMultiplierBondnonlinear", NULL, NULL, "dampingNormalForceBondnonlinear",
    "dampingTangentialForceBondnonlinear", "dampingNormalTorqueBondnonlinear",
    "dampingTangentialTorqueBondnonlinear", "maxDistanceBondnonlinear", "maxSigmaBondnonlinear",
    "maxTauBondnonlinear", "ratioTensionCompressionBondnonlinear", "createDistanceBondnonlinear"};

bool on_off(Error *error, const char *kw, const char *v)
{
  if (strcmp(v, "on") == 0 || strcmp(v, "yes") == 0) return true;
  if (strcmp(v, "off") == 0 || strcmp(v, "no") == 0) return false;
  std::string msg = std::string("pair gran/mi355x: expecting on/off after ") + kw;
  error->all(FLERR, msg.c_str());
  return false;
}

}  // namespace

PairGranMI355X::PairGranMI355X(LAMMPS *lmp) :
  PairGran(lmp),
  normal_(DEM_NORMAL_HERTZ), rolling_(DEM_ROLL_OFF), cohesion_(DEM_COHESION_OFF),
  limit_force_(0), tangential_damping_(1), torsion_torque_(0), mixed_(0), krs_(0.0),
  Yeff_(NULL), Geff_(NULL), betaeff_(NULL), coeffFrict_(NULL), coeffRollFrict_(NULL),
  coeffRollVisc_(NULL), coeffRestLog_(NULL), coeffRollStiffness_(0.0)
{
  memset(&bond_, 0, sizeof(bond_));
  bond_.damping = 1;
  // the contact law's precision is a property of the device run, not of
  // the script: LIGGGHTS_MI355X_PRECISION=mixed selects the FP32 law
  const char *prec = getenv("LIGGGHTS_MI355X_PRECISION");
  mixed_ = prec && strcmp(prec, "mixed") == 0;
}

PairGranMI355X::~PairGranMI355X() {}

/* pair_style gran model <normal> tangential <t> [cohesion <c>]
   [rolling_friction <r>] [surface default] [<setting> <value> ...]
   (pair_gran_proxy.cpp settings, contact_models.h registerSettings) */
void PairGranMI355X::settings(int narg, char **arg)
{
  bool have_model = false, have_tangential = false;
  int iarg = 0;
  while (iarg < narg) {
    const char *kw = arg[iarg];
    if (iarg + 1 >= narg) error->all(FLERR, "pair gran/mi355x: missing value after keyword");
    const char *v = arg[iarg + 1];
    if (strcmp(kw, "model") == 0) {
      if (strcmp(v, "hertz") == 0) normal_ = DEM_NORMAL_HERTZ, limit_force_ = 0;
      else if (strcmp(v, "hysteretic/nonlinear1") == 0) normal_ = DEM_NORMAL_HYST1, limit_force_ = 1;
      else if (strcmp(v, "hysteretic/nonlinear2") == 0) normal_ = DEM_NORMAL_HYST2, limit_force_ = 1;
      else error->all(FLERR, "pair gran/mi355x: normal model not on the device (hertz, hysteretic/nonlinear1/2)");
      have_model = true;
    } else if (strcmp(kw, "tangential") == 0) {
      if (strcmp(v, "history") != 0) error->all(FLERR, "pair gran/mi355x: tangential history only");
      have_tangential = true;
    } else if (strcmp(kw, "rolling_friction") == 0) {
      if (strcmp(v, "off") == 0) rolling_ = DEM_ROLL_OFF;
      else if (strcmp(v, "epsd") == 0) rolling_ = DEM_ROLL_EPSD;
      else if (strcmp(v, "epsd2") == 0) rolling_ = DEM_ROLL_EPSD2;
      else if (strcmp(v, "epsd3") == 0) rolling_ = DEM_ROLL_EPSD3;
      else if (strcmp(v, "cdtnonlinear2") == 0) rolling_ = DEM_ROLL_CDTNONLINEAR2;
      else error->all(FLERR, "pair gran/mi355x: rolling model not on the device");
    } else if (strcmp(kw, "cohesion") == 0) {
      if (strcmp(v, "off") == 0) cohesion_ = DEM_COHESION_OFF;
      else if (strcmp(v, "bond") == 0) cohesion_ = DEM_COHESION_BOND;
      else if (strcmp(v, "bond/nonlinear") == 0) cohesion_ = DEM_COHESION_BOND_NONLINEAR;
      else error->all(FLERR, "pair gran/mi355x: cohesion model not on the device (bond, bond/nonlinear)");
    } else if (strcmp(kw, "surface") == 0) {
      if (strcmp(v, "default") != 0) error->all(FLERR, "pair gran/mi355x: surface default only");
    } else if (strcmp(kw, "limitForce") == 0) {
      limit_force_ = on_off(error, kw, v);
    } else if (strcmp(kw, "tangential_damping") == 0) {
      tangential_damping_ = on_off(error, kw, v);
    } else if (strcmp(kw, "torsionTorque") == 0) {
      torsion_torque_ = on_off(error, kw, v);
    } else if (strcmp(kw, "stressBreak") == 0) {
      bond_.stress_break = on_off(error, kw, v);
    } else if (strcmp(kw, "createBondAlways") == 0) {
      bond_.create_always = on_off(error, kw, v);
    } else if (strcmp(kw, "dampingBond") == 0) {
      bond_.damping = on_off(error, kw, v);
    } else if (strcmp(kw, "dampingBondSmooth") == 0) {
      bond_.damping_smooth = on_off(error, kw, v);
      if (bond_.damping_smooth) bond_.damping = 1;
    } else if (strcmp(kw, "ratioTensionCompressionBond") == 0) {
      bond_.ratio_tc = on_off(error, kw, v);
    } else {
      std::string msg = std::string("pair gran/mi355x: setting not on the device: ") + kw;
      error->all(FLERR, msg.c_str());
    }
    iarg += 2;
  }
  if (!have_model || !have_tangential)
    error->all(FLERR, "pair gran/mi355x: needs 'model <normal> tangential history'");
  register_history();
}

void PairGranMI355X::register_history()
{
  history_arg.clear();
  dnum_pairgran = 0;
  history = false;
  if (normal_ != DEM_NORMAL_HERTZ)
    for (unsigned k = 0; k < sizeof(kHysteretic) / sizeof(kHysteretic[0]); k++)
      add_history_value(kHysteretic[k].name, kHysteretic[k].newton);
  if (cohesion_ == DEM_COHESION_BOND)
    for (unsigned k = 0; k < sizeof(kBond) / sizeof(kBond[0]); k++)
      add_history_value(kBond[k].name, kBond[k].newton);
  if (cohesion_ == DEM_COHESION_BOND_NONLINEAR)
    for (unsigned k = 0; k < sizeof(kBondNonlinear) / sizeof(kBondNonlinear[0]); k++)
      add_history_value(kBondNonlinear[k].name, kBondNonlinear[k].newton);
  for (unsigned k = 0; k < 3; k++) add_history_value(kShear[k].name, kShear[k].newton);
  if (rolling_ != DEM_ROLL_OFF)
    for (unsigned k = 0; k < 3; k++) add_history_value(kRoll[k].name, kRoll[k].newton);
}

/* The effective tables come from the reference's own creators
   (global_properties.cpp:428-584), connected as the models connect them. */
void PairGranMI355X::init_granular()
{
  PropertyRegistry &reg = force->registry;
  const char *caller = "pair gran/mi355x";
  reg.registerProperty("Yeff", &MODEL_PARAMS::createYeff);
  reg.registerProperty("Geff", &MODEL_PARAMS::createGeff);
  reg.registerProperty("betaeff", &MODEL_PARAMS::createBetaEff);
  reg.registerProperty("coeffFrict", &MODEL_PARAMS::createCoeffFrict);
  reg.connect("Yeff", Yeff_, caller);
  reg.connect("Geff", Geff_, caller);
  reg.connect("betaeff", betaeff_, caller);
  reg.connect("coeffFrict", coeffFrict_, caller);
  if (rolling_ == DEM_ROLL_EPSD || rolling_ == DEM_ROLL_EPSD2 || rolling_ == DEM_ROLL_EPSD3
      || rolling_ == DEM_ROLL_CDTNONLINEAR2) {
    reg.registerProperty("coeffRollFrict", &MODEL_PARAMS::createCoeffRollFrict);
    reg.connect("coeffRollFrict", coeffRollFrict_, caller);
  }
  if (rolling_ == DEM_ROLL_EPSD || rolling_ == DEM_ROLL_EPSD3) {
    reg.registerProperty("coeffRollVisc", &MODEL_PARAMS::createCoeffRollVisc);
    reg.connect("coeffRollVisc", coeffRollVisc_, caller);
  }
  if (rolling_ == DEM_ROLL_EPSD3) {
    reg.registerProperty("coeffRollStiffness", &MODEL_PARAMS::createRollingStiffness);
    reg.connect("coeffRollStiffness", coeffRollStiffness_, caller);
    krs_ = coeffRollStiffness_;
  }
  if (normal_ != DEM_NORMAL_HERTZ) {
    reg.registerProperty("CoeffRestLog", &MODEL_PARAMS::createCoeffRestLog);
    reg.connect("CoeffRestLog", coeffRestLog_, caller);
  }
}

void PairGranMI355X::compute_force(int, int)
{
  error->all(FLERR, "pair gran/mi355x computes on the device: use run_style verlet/mi355x "
                    "(implied by -sf mi355x)");
}

bool PairGranMI355X::contact_match(const std::string mtype, const std::string model)
{
  if (mtype == "model")
    return (model == "hertz" && normal_ == DEM_NORMAL_HERTZ)
           || (model == "hysteretic/nonlinear1" && normal_ == DEM_NORMAL_HYST1)
           || (model == "hysteretic/nonlinear2" && normal_ == DEM_NORMAL_HYST2);
  if (mtype == "tangential") return model == "history";
  if (mtype == "cohesion")
    return (model == "off" && cohesion_ == DEM_COHESION_OFF) || (model == "bond" && cohesion_ == DEM_COHESION_BOND)
           || (model == "bond/nonlinear" && cohesion_ == DEM_COHESION_BOND_NONLINEAR);
  if (mtype == "rolling_friction")
    return (model == "off" && rolling_ == DEM_ROLL_OFF) || (model == "epsd" && rolling_ == DEM_ROLL_EPSD)
           || (model == "epsd2" && rolling_ == DEM_ROLL_EPSD2) || (model == "epsd3" && rolling_ == DEM_ROLL_EPSD3)
           || (model == "cdtnonlinear2" && rolling_ == DEM_ROLL_CDTNONLINEAR2);
  if (mtype == "surface") return model == "default";
  return false;
}

int64_t PairGranMI355X::hashcode()
{
  // distinct per chain, as the reference's generated model hash is
  return 0x4d49333535580000LL + normal_ + 16 * rolling_ + 256 * cohesion_;
}

double PairGranMI355X::stressStrainExponent()
{
  return normal_ == DEM_NORMAL_HERTZ ? 1.5 : 1.0;
}

int PairGranMI355X::wall_normal_model_of(const char *name) const
{
  if (strcmp(name, "hertz") == 0) return DEM_NORMAL_HERTZ;
  if (strcmp(name, "hysteretic/nonlinear1") == 0) return DEM_NORMAL_HYST1;
  if (strcmp(name, "hysteretic/nonlinear2") == 0) return DEM_NORMAL_HYST2;
  return -1;
}

void PairGranMI355X::fill_params(dem_params *p) const
{
  p->normal_model = normal_;
  p->rolling_model = rolling_;
  p->cohesion_model = cohesion_;
  p->limit_force = limit_force_;
  p->tangential_damping = tangential_damping_;
  p->torsion_torque = torsion_torque_;
  p->precision = mixed_ ? DEM_MIXED : DEM_FP64;
}

std::vector<double> PairGranMI355X::flat(double **m) const
{
  const int w = atom->ntypes + 1;
  std::vector<double> out(w * w, 0.0);
  if (m)
    for (int i = 1; i < w; i++)
      for (int j = 1; j < w; j++) out[i * w + j] = m[i][j];
  return out;
}

// a peratomtypepair property/global table, (ntypes+1)^2 with row/col 0 unused
std::vector<double> PairGranMI355X::pair_table(const char *name, bool required) const
{
  const int nt = atom->ntypes, w = nt + 1;
  std::vector<double> out(w * w, 0.0);
  FixPropertyGlobal *fp = static_cast<FixPropertyGlobal *>(
      modify->find_fix_property(name, "property/global", "peratomtypepair", nt, nt, "pair gran/mi355x", required));
  if (!fp) return out;
  for (int i = 1; i < w; i++)
    for (int j = 1; j < w; j++) out[i * w + j] = fp->compute_array(i - 1, j - 1);
  return out;
}

int PairGranMI355X::upload_materials(dem_ctx *ctx) const
{
  const int nt = atom->ntypes;
  std::vector<double> Y = flat(Yeff_), G = flat(Geff_), B = flat(betaeff_), mu = flat(coeffFrict_);
  std::vector<double> mur = flat(coeffRollFrict_), etar = flat(coeffRollVisc_);
  int rc = dem_set_materials(ctx, nt, Y.data(), G.data(), B.data(), mu.data(),
                             coeffRollFrict_ ? mur.data() : NULL, coeffRollVisc_ ? etar.data() : NULL, krs_);
  if (rc || normal_ == DEM_NORMAL_HERTZ) return rc;
  std::vector<double> crl = flat(coeffRestLog_);
  std::vector<double> tabs;
  for (int k = 0; k < DEM_NHYST; k++) {
    // LoadingStiffness is required; the others default to the model's
    // own defaults when absent (normal_model_hysteretic_nonlinear1.h:105-133)
    std::vector<double> t = pair_table(kHystProps[k], k == 0);
    tabs.insert(tabs.end(), t.begin(), t.end());
  }
  return dem_set_hysteretic(ctx, crl.data(), tabs.data());
}

int PairGranMI355X::upload_bond(dem_ctx *ctx) const
{
  if (cohesion_ == DEM_COHESION_OFF) return 0;
  const bool nl = cohesion_ == DEM_COHESION_BOND_NONLINEAR;
  std::vector<double> tabs;
  for (int k = 0; k < DEM_NBOND; k++) {
    const char *name = kBondProps[k];
    if (nl && k < 12) name = kBondNonlinearProps[k];
    std::vector<double> t;
    if (name) t = pair_table(name, false);
    else t.assign((atom->ntypes + 1) * (atom->ntypes + 1), 0.0);
    tabs.insert(tabs.end(), t.begin(), t.end());
  }
  dem_bond_settings bs = bond_;
  FixPropertyGlobal *fts = static_cast<FixPropertyGlobal *>(modify->find_fix_property(
      nl ? "tsCreateBondnonlinear" : "tsCreateBond", "property/global", "scalar", 0, 0, "pair gran/mi355x", false));
  bs.ts_create = fts ? fts->compute_scalar() : 0.0;
  return dem_set_bond(ctx, tabs.data(), &bs);
}
