/* run_style verlet/mi355x: the device owns the step between host sync
 * points (setup, output, end-of-step host fixes, end of run). */
#include "verlet_mi355x.h"

#include "atom.h"
#include "comm.h"
#include "domain.h"
#include "error.h"
#include "fix.h"
#include "fix_contact_history.h"
#include "fix_force_ops_mi355x.h"
#include "fix_gravity.h"
#include "fix_massflow_mesh_mi355x.h"
#include "fix_mesh_surface_mi355x.h"
#include "fix_move_mesh_mi355x.h"
#include "fix_move_mi355x.h"
#include "fix_nve_sphere_mi355x.h"
#include "fix_wall_gran_mi355x.h"
#include "force.h"
#include "group.h"
#include "modify.h"
#include "neighbor.h"
#include "output.h"
#include "pair_gran_mi355x.h"
#include "timer.h"
#include "update.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <string>

using namespace LAMMPS_NS;
using namespace FixConst;

namespace {

// host fixes that are passive during a run (storage, lists, checks), or
// whose work the device does from the values the styles hand over
bool device_or_passive(Fix *f)
{
  static const char *names[] = {"gravity", "freeze", "contacthistory", "contacthistory/mesh",
                                "property/global", "property/atom", "neighlist/mesh",
                                "check/timestep/gran", "print", "ave/time", NULL};
  if (dynamic_cast<FixWallGranMI355X *>(f) || dynamic_cast<FixMeshSurfaceMI355X *>(f)
      || dynamic_cast<FixMoveMeshMI355X *>(f) || dynamic_cast<FixNVESphereMI355X *>(f)
      || dynamic_cast<FixSetForceMI355X *>(f) || dynamic_cast<FixAveForceMI355X *>(f)
      || dynamic_cast<FixMoveMI355X *>(f) || dynamic_cast<FixMassflowMeshMI355X *>(f))
    return true;
  for (int k = 0; names[k]; k++)
    if (strcmp(f->style, names[k]) == 0) return true;
  return false;
}

// group bit of a group mask (Group::bitmask is 1 << igroup)
int group_bit(int groupbit)
{
  int b = 0;
  while (b < 31 && !(groupbit & (1 << b))) b++;
  return b;
}

}  // namespace

VerletMI355X::VerletMI355X(LAMMPS *lmp, int narg, char **arg) :
  Verlet(lmp, narg, arg), ctx_(NULL), pair_(NULL), last_download_(-1), have_device_history_(false),
  nplanes_(0), nmeshes_(0)
{
}

VerletMI355X::~VerletMI355X()
{
  if (ctx_) dem_ctx_destroy(ctx_);
}

void VerletMI355X::check(int rc)
{
  if (rc) error->one(FLERR, dem_last_error(ctx_));
}

void VerletMI355X::init()
{
  Verlet::init();
  if (comm->nprocs != 1) error->all(FLERR, "verlet/mi355x runs one process per GPU: one MPI rank");
  pair_ = dynamic_cast<PairGranMI355X *>(force->pair);
  if (!pair_) error->all(FLERR, "verlet/mi355x needs pair_style gran (run with -sf mi355x)");
  if (atom->rmass == NULL || atom->radius == NULL || atom->omega == NULL)
    error->all(FLERR, "verlet/mi355x needs atom_style granular");
  for (int i = 0; i < modify->nfix; i++) {
    Fix *f = modify->fix[i];
    if (device_or_passive(f)) continue;
    const int mask = modify->fmask[i];
    if (mask & (INITIAL_INTEGRATE | POST_INTEGRATE | PRE_FORCE | POST_FORCE | FINAL_INTEGRATE)) {
      std::string msg = std::string("verlet/mi355x: fix ") + f->style
                        + " acts inside the step on the host and has no device form; run without -sf mi355x";
      error->all(FLERR, msg.c_str());
    }
  }
}

/* The context mirrors the script: box, neighbour settings, the pair
   chain, then walls in fix order, gravity and freeze. */
void VerletMI355X::create_context()
{
  dem_params p;
  memset(&p, 0, sizeof(p));
  p.dt = update->dt;
  for (int d = 0; d < 3; d++) {
    p.boxlo[d] = domain->boxlo[d];
    p.boxhi[d] = domain->boxhi[d];
    p.periodic[d] = domain->periodicity[d];
  }
  p.skin = neighbor->skin;
  p.cdf = neighbor->contactDistanceFactor;
  p.neigh_every = neighbor->every;
  p.neigh_delay = neighbor->delay;
  p.neigh_check = neighbor->dist_check;
  p.neigh_one = neighbor->oneatom;
  p.device = 0;
  pair_->fill_params(&p);
  p.wall_normal_model = p.normal_model;
  p.wall_limit_force = p.limit_force;
  for (int i = 0; i < modify->nfix; i++)
    if (FixWallGranMI355X *w = dynamic_cast<FixWallGranMI355X *>(modify->fix[i])) {
      p.wall_normal_model = w->normal_model();
      p.wall_limit_force = w->limit_force();
    }
  ctx_ = dem_ctx_create(&p);
  if (!ctx_) error->one(FLERR, dem_last_error(NULL));
  nplanes_ = nmeshes_ = 0;
  for (int i = 0; i < modify->nfix; i++)
    if (FixWallGranMI355X *w = dynamic_cast<FixWallGranMI355X *>(modify->fix[i])) {
      check(w->register_with(ctx_));
      if (w->is_mesh()) nmeshes_ += w->n_meshes();
      else nplanes_++;
    }
  double g[3] = {0.0, 0.0, 0.0};
  if (FixGravity *fg = static_cast<FixGravity *>(modify->find_fix_style("gravity", 0))) {
    fg->get_gravity(g);
    check(dem_set_gravity(ctx_, g[0], g[1], g[2]));
  }
  check(dem_set_freeze(ctx_, modify->find_fix_style("freeze", 0) != NULL));
  check(pair_->upload_materials(ctx_));
}

void VerletMI355X::upload()
{
  const int n = atom->nlocal;
  std::vector<int> frz(n, 0);
  const int fb = pair_->freeze_group_bit();
  for (int i = 0; i < n; i++) frz[i] = (atom->mask[i] & fb) != 0;
  check(dem_upload_state(ctx_, n, atom->x[0], atom->v[0], atom->omega[0], atom->radius, atom->rmass,
                         atom->type, atom->tag, frz.data()));
  check(pair_->upload_bond(ctx_));
}

/* Contact history into the context: the device's own from the previous
   run command, or on the first run the host's FixContactHistory pages
   (read_restart, fix_contact_history.cpp:588-656). */
void VerletMI355X::upload_history()
{
  if (have_device_history_) {
    if (!hist_i_.empty())
      check(dem_set_history(ctx_, (int64_t)hist_i_.size(), hist_i_.data(), hist_j_.data(), hist_.data()));
    // primitive-wall history, by tag into the new upload order
    const int n = atom->nlocal;
    const size_t per = wall_tag_.empty() ? 0 : wall_hist_.size() / wall_tag_.size();
    if (per) {
      std::map<int, size_t> where;
      for (size_t k = 0; k < wall_tag_.size(); k++) where[wall_tag_[k]] = k;
      std::vector<double> h((size_t)n * per, 0.0);
      for (int i = 0; i < n; i++) {
        std::map<int, size_t>::const_iterator it = where.find(atom->tag[i]);
        if (it != where.end())
          std::copy(wall_hist_.begin() + it->second * per, wall_hist_.begin() + (it->second + 1) * per,
                    h.begin() + (size_t)i * per);
      }
      check(dem_set_wall_history(ctx_, h.data()));
    }
    for (size_t m = 0; m < mesh_slots_.size(); m++) {
      const MeshSlots &ms = mesh_slots_[m];
      if (!ms.tag.empty())
        check(dem_set_mesh_contacts(ctx_, (int)m, (int64_t)ms.tag.size(), ms.tag.data(), ms.tri.data(),
                                    ms.slot.data(), ms.hist.data()));
    }
    return;
  }
  FixContactHistory *fh = pair_->get_fix_history();
  if (!fh) return;
  const int dnum = fh->get_dnum();
  if (dnum != dem_history_size(ctx_)) error->all(FLERR, "verlet/mi355x: contact history layout differs");
  std::vector<int> ti, tj;
  std::vector<double> h;
  for (int i = 0; i < atom->nlocal; i++)
    for (int k = 0; k < fh->n_partner(i); k++) {
      ti.push_back(atom->tag[i]);
      tj.push_back(fh->partner(i, k));
      const double *src = fh->contacthistory(i, k);
      h.insert(h.end(), src, src + dnum);
    }
  if (!ti.empty()) check(dem_set_history(ctx_, (int64_t)ti.size(), ti.data(), tj.data(), h.data()));
}

// members of a fix's group as dem_set_group takes them; -1 for group all
int VerletMI355X::device_group(Fix *f)
{
  if (f->igroup == 0) return -1;
  const int n = atom->nlocal, bit = group_bit(f->groupbit);
  std::vector<int> members(n);
  for (int k = 0; k < n; k++) members[k] = (atom->mask[k] & f->groupbit) != 0;
  check(dem_set_group(ctx_, bit, members.data()));
  return bit;
}

/* The per-atom fixes on the device, after the upload (groups are per-atom
   bits): fix move (before the integrator, as in scripts that define it
   first), nve/sphere on its group, the group force fixes in fix order, and
   the massflow counters. */
void VerletMI355X::register_force_ops()
{
  int nnve = 0;
  for (int i = 0; i < modify->nfix; i++) {
    Fix *f = modify->fix[i];
    if (FixMoveMI355X *mv = dynamic_cast<FixMoveMI355X *>(f)) {
      const int bit = device_group(f);
      int id = 0;
      check(mv->register_with(ctx_, bit < 0 ? group_bit(f->groupbit) : bit, &id));
    } else if (dynamic_cast<FixNVESphereMI355X *>(f)) {
      if (++nnve > 1) error->all(FLERR, "verlet/mi355x: one fix nve/sphere on the device");
      check(dem_set_integrate_group(ctx_, device_group(f)));
    } else if (FixMassflowMeshMI355X *mf = dynamic_cast<FixMassflowMeshMI355X *>(f)) {
      check(mf->register_with(ctx_, device_group(f)));
    }
  }
  for (int i = 0; i < modify->nfix; i++) {
    Fix *f = modify->fix[i];
    FixSetForceMI355X *sf = dynamic_cast<FixSetForceMI355X *>(f);
    FixAveForceMI355X *af = dynamic_cast<FixAveForceMI355X *>(f);
    if (!sf && !af) continue;
    const int n = atom->nlocal, bit = group_bit(f->groupbit);
    std::vector<int> members(n);
    for (int k = 0; k < n; k++) members[k] = (atom->mask[k] & f->groupbit) != 0;
    check(dem_set_group(ctx_, bit, members.data()));
    int id = 0;
    check(sf ? sf->register_with(ctx_, bit, &id) : af->register_with(ctx_, bit, &id));
  }
}

void VerletMI355X::download()
{
  check(dem_download(ctx_, atom->x[0], atom->v[0], atom->omega[0], atom->f[0], atom->torque[0]));
  last_download_ = update->ntimestep;
}

void VerletMI355X::save_device_history()
{
  int64_t count = 0;
  check(dem_get_pairs(ctx_, &count, NULL, NULL, NULL, NULL, NULL));
  const int dnum = dem_history_size(ctx_);
  std::vector<int> ti(count), tj(count), sh(count), fl(count);
  std::vector<double> h((size_t)count * dnum);
  check(dem_get_pairs(ctx_, &count, ti.data(), tj.data(), sh.data(), h.data(), fl.data()));
  hist_i_.clear();
  hist_j_.clear();
  hist_.clear();
  for (int64_t e = 0; e < count; e++) {
    if (!fl[e]) continue;
    hist_i_.push_back(ti[e]);
    hist_j_.push_back(tj[e]);
    hist_.insert(hist_.end(), h.begin() + e * dnum, h.begin() + (e + 1) * dnum);
  }
  // primitive-wall history (upload order) keyed by tag
  const int n = atom->nlocal;
  const size_t per = (size_t)dem_wall_history_size(ctx_) * nplanes_;
  wall_tag_.assign(atom->tag, atom->tag + n);
  wall_hist_.assign((size_t)n * per, 0.0);
  if (per) check(dem_get_wall_history(ctx_, wall_hist_.data()));
  // mesh contact slots, as (tag, triangle, slot, history) records
  mesh_slots_.assign(nmeshes_, MeshSlots());
  const int dw = dem_wall_history_size(ctx_);
  for (int m = 0; m < nmeshes_; m++) {
    MeshSlots &ms = mesh_slots_[m];
    int64_t c = 0;
    check(dem_get_mesh_contacts(ctx_, m, &c, NULL, NULL, NULL, NULL));
    ms.tag.resize(c);
    ms.tri.resize(c);
    ms.slot.resize(c);
    ms.hist.resize((size_t)c * dw);
    if (c) check(dem_get_mesh_contacts(ctx_, m, &c, ms.tag.data(), ms.tri.data(), ms.slot.data(), ms.hist.data()));
  }
  have_device_history_ = true;
}

/* Verlet::setup (verlet.cpp:134-199) with the force computation on the
   device: the host keeps domain, comm and neighbour bookkeeping. */
void VerletMI355X::setup()
{
  update->setupflag = 1;
  atom->setup();
  modify->setup_pre_exchange();
  domain->pbc();
  domain->reset_box();
  comm->setup();
  if (neighbor->style) neighbor->setup_bins();
  comm->exchange();
  if (atom->sortfreq > 0) atom->sort();
  comm->borders();
  modify->setup_pre_neighbor();
  neighbor->build();
  neighbor->ncalls = 0;
  ev_set(update->ntimestep);
  force_clear();

  if (!ctx_) create_context();
  upload();
  upload_history();
  register_force_ops();
  check(dem_setup(ctx_));
  download();

  // host fixes that are not on the device keep their set-up call (an
  // end-of-step fix samples the set-up state); the device fixes' host
  // forms would add their forces a second time
  for (int i = 0; i < modify->nfix; i++) {
    Fix *f = modify->fix[i];
    if (dynamic_cast<FixWallGranMI355X *>(f) || dynamic_cast<FixNVESphereMI355X *>(f)
        || dynamic_cast<FixSetForceMI355X *>(f) || dynamic_cast<FixAveForceMI355X *>(f)
        || dynamic_cast<FixMoveMeshMI355X *>(f) || dynamic_cast<FixMoveMI355X *>(f)
        || dynamic_cast<FixMassflowMeshMI355X *>(f) || strcmp(f->style, "gravity") == 0
        || strcmp(f->style, "freeze") == 0)
      continue;
    if (modify->fmask[i] & END_OF_STEP) f->setup(vflag);
  }
  output->setup();
  update->setupflag = 0;
}

// the next step at which the host needs the state: output, or an
// end-of-step host fix
bigint VerletMI355X::next_host_step(bigint stop) const
{
  bigint next = stop < output->next ? stop : output->next;
  for (int i = 0; i < modify->nfix; i++) {
    if (!(modify->fmask[i] & END_OF_STEP)) continue;
    const int every = modify->fix[i]->nevery;
    if (every <= 0) continue;
    const bigint due = (update->ntimestep / every + 1) * every;
    if (due < next) next = due;
  }
  return next;
}

void VerletMI355X::run(int n)
{
  const bigint stop = update->ntimestep + n;
  while (update->ntimestep < stop) {
    const bigint next = next_host_step(stop);
    timer->stamp();
    check(dem_step(ctx_, (int)(next - update->ntimestep)));
    update->ntimestep = next;
    timer->stamp(TIME_PAIR);
    if (next == output->next || modify->n_end_of_step) {
      download();
      if (modify->n_end_of_step) modify->end_of_step();
      timer->stamp(TIME_MODIFY);
    }
    if (next == output->next) {
      output->write(next);
      timer->stamp(TIME_OUTPUT);
    }
  }
}

/* Verlet::cleanup plus the device's end of run: the state comes back to
   the host arrays, the run time bookkeeping advances (dem_cleanup is
   Update::update_time on the device), and the contact history is kept for
   the next run command. */
void VerletMI355X::cleanup()
{
  if (last_download_ != update->ntimestep) download();
  check(dem_cleanup(ctx_));
  save_device_history();
  Verlet::cleanup();
}
