/* run_style verlet/mi355x -- Verlet::run with the particle hot path on the
 * MI355X (liggghts_dem C-ABI).
 *
 * With `-sf mi355x` the default `run_style verlet` resolves to this style
 * (update.cpp:373-385).  It owns the timestep: setup() builds the device
 * context from the script's pair style, walls, meshes, gravity, freeze and
 * group force fixes, uploads the atoms with their contact history and runs
 * the set-up force pass; run(n) advances the device up to each step at
 * which output or an end-of-step host fix needs the state, downloads it and
 * hands over to Output/Modify as Verlet::run does (verlet.cpp:264-390).
 * A fix that must act inside the step on the host has no place here: init()
 * refuses it by name, and plain `verlet` (no -sf) keeps running it.
 */
#ifdef INTEGRATE_CLASS

IntegrateStyle(verlet/mi355x,VerletMI355X)

#else

#ifndef LMP_VERLET_MI355X_H
#define LMP_VERLET_MI355X_H

#include "liggghts_dem.h"
#include "verlet.h"

#include <vector>

namespace LAMMPS_NS {

class VerletMI355X : public Verlet {
 public:
  VerletMI355X(class LAMMPS *, int, char **);
  virtual ~VerletMI355X();
  virtual void init();
  virtual void setup();
  virtual void run(int);
  virtual void cleanup();

 private:
  dem_ctx *ctx_;
  class PairGranMI355X *pair_;
  bigint last_download_;
  // contact history carried from one run command to the next on the
  // device, keyed by tags (atoms may be re-sorted between runs)
  std::vector<int> hist_i_, hist_j_;
  std::vector<double> hist_;
  std::vector<int> wall_tag_;          // primitive-wall history per atom tag
  std::vector<double> wall_hist_;
  struct MeshSlots {                   // mesh contact slots of one mesh
    std::vector<int> tag, tri, slot;
    std::vector<double> hist;
  };
  std::vector<MeshSlots> mesh_slots_;
  bool have_device_history_;
  int nplanes_, nmeshes_;

  void check(int rc);
  void create_context();
  void upload();
  void upload_history();
  void register_force_ops();
  int device_group(class Fix *f);
  void download();
  void save_device_history();
  bigint next_host_step(bigint stop) const;
};

}

#endif
#endif
