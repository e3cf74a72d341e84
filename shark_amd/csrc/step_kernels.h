// step_kernels.h — argument block and launchers of the per-step kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace shark {

// Scalars that live on the device so consecutive steps need no host round trip.
struct DevScalars {
  double dt;            // dt of the step being taken (computed at the end of the previous one)
  double time;          // simulated time before the step
  double tend;          // run stops once time >= tend (flag_continue)
  double dt_last;       // dt of the last completed step
  unsigned long long dtbits;  // running min of next-step dt candidates (ordered bit pattern)
  long long nstep;      // completed steps
  int active;           // 0 once time >= tend: every kernel returns immediately
  int pad;
  // the planet at the step's time [0..1] and at time + dt [2..3] (the next step's Courant
  // candidate): cartesian disk (disk/setup.f90:292-301) Mplanet x_p / r_p^3 and Mplanet y_p / r_p^3,
  // polar disk (disk_polar/setup.f90:331-334) cos(theta_p) Rplanet and sin(theta_p) Rplanet
  double planet[4];
};

// Geometry of a separable grid (every reference grid is: gridinit builds dx, surf, vol and
// radii_c from ix alone, src/grid/grid.f90:404-455,480-496): one entry per column ix.
struct ColGeom {
  const double* dx1;    // dx(:,1)
  const double* dx2;    // dx(:,2)
  const double* surf1;  // surf(:,1)
  const double* surf2;  // surf(:,2)
  const double* vol;
  const double* rc;     // radii_c
  const double* pos1;   // position(:,1)
  const double* pos2;   // position(:,2) per row iy (cartesian 2D: separable too)
};

struct StepArgs {
  double* U;            // state, (nvar, ncells)
  double* B;            // post-hydro state of the current step, (nvar, ncells)
  const double* cs;     // (ncells)
  ColGeom g;
  const double* sdust;  // (ndust, ncells) or nullptr when sdust_uniform
  const double* condinit;   // (nvar, ncells), disk in-loop hook
  const double* fargo_vk;   // fargo_velocity per column
  double* Rold;             // (ndust, ncells) dust densities before growth (growth in the step)
  // (4, ncells) gas face sound speeds for the HLLC dust solver with adiabatic gas, written by the
  // gas hydro pass: [0]/[1] left/right state of the x face on the cell's left, [2]/[3] of the y
  // face below it; nullptr when the dust solver does not read them
  double* fcs;
  DevScalars* sc;
  int nx, ny, nd, nxm, nym, nc, nvar, ndim;
  int geom, solver, solverdust, slope, force_kick, force_kind, tstop_kind, inloop_kind;
  int bc_kind, drag, back_reaction, fargo, disk_bc_mode, sdust_uniform;
  int strip_h;          // rows per dust hydro work item
  int ngroups, nstrips; // hydro decomposition (dust)
  int strip_g, nstrips_g;  // rows per gas hydro work item, and the gas strips
  int post_tw;          // cells per post row segment (post_tw_ok)
  int xcd_order;        // 1: XCD-aware block order in the hydro and post launches
  double gamma, cfl, smallr, dust_ratio_min;
  double q_shear, omega_shear, box_l, rad0, eta_stream, rho_init, hoverr, rhograin;
  double mstar, r_relax, n_rel, smooth_r;
  double K[64];
  double sdust_k[64];   // per-species grain size when uniform over cells
  double rho_l, rho_r, p_l, p_r;   // Sod tube states
  int shock_dir;        // Sod tube direction
  int iso;              // 1: isothermal (setup's cs); 0: adiabatic (P from E, cs = sqrt(gamma P / rho))
  double smallp;
  double eps_k1;        // sqrt(pi gamma / 8) * rhograin: the Epstein prefactor (disk tstop)
  double s2pi;          // sqrt(2 pi)
  // cartesian disk (src/setups/disk): second box length, planet, initial disk, buffer zones,
  // alpha viscosity (eta_visc = alpha_visc H cs), dust-to-gas ratio of the buffers
  double box_l_y, mplanet, rplanet, sigma_r0, r0_disk, frac_r_disk, buffer_in, buffer_out;
  double alpha_visc, dust2gas;
  int viscosity;
  int accretion;        // polar disk inflow through the outer boundary (boundaries.f90:63-94)
  double m_acc, phi_mom, l_out;
  // Collapse_1D (spherical 1D grid): 1D self-gravity and the barotropic EOS
  int gravity, eos_baro;
  double l_soft, m_cent, t0_cloud, mu_gas, unit_v, unit_nh;
  const double* dvol;   // (ncells) part of each cell between its centre and its outer face
  double* mc;           // (ncells) mass enclosed within each cell centre, this step (k_mtot)
  // polar-disk planet (disk_polar/setup.f90:323-378): softening length, the grid's position
  // (2, ncells) for its potential, and the gradient of the potential at the step's time and at
  // time + dt (4, ncells), written by k_planet_grad; nullptr without a planet
  double a_smooth_pl;
  const double* posxy;
  double* pgrad;
  // polar-disk self-gravity (disk_polar/setup.f90:300-322; replaced by the planet's potential when
  // there is a planet): the potential per cell, and cos(phi_p - phi_loc) per (source row, target
  // row) of the active rows, computed on the host with the host's cos; nullptr without it
  int self_gravity;
  double* phi_sg;
  const double* sgcos;
  // explicit drag in the force (drag = .false. with force_kick; SI setup.f90:355-362): every
  // fluid's force of the step, (3 (1 + ndust), ncells), from k_explicit_force; nullptr otherwise
  double* fexp;
};

// launch shapes: every shape gives the same results bit for bit, only the speed differs
bool post_tw_ok(int nd, int tw);
int post_tw_default(int nd);
int post_tw_candidates(int nd, int* out);   // returns the count (<= 4)

void launch_bc(const StepArgs& a, hipStream_t s);
void launch_hydro(const StepArgs& a, hipStream_t s);
void launch_hydro_fluid(const StepArgs& a, bool gas, hipStream_t s);   // the gas or the dust launch
void launch_post(const StepArgs& a, hipStream_t s);
// the post pass split around a growth step: part 1 remap + drag, part 2 in-loop hook + Courant
void launch_post_part(const StepArgs& a, int part, hipStream_t s);
void launch_courant_init(const StepArgs& a, hipStream_t s);
void launch_finalize(const StepArgs& a, hipStream_t s);
void launch_set_dt(const StepArgs& a, hipStream_t s);
void launch_arm(const StepArgs& a, double tend, hipStream_t s);
void launch_ctoprim(const StepArgs& a, double* q, hipStream_t s);

}  // namespace shark
