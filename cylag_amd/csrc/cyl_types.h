// Device-side views of the mesh, the Eulerian fields and the particle pool.
//
// Layout on HBM (one GPU holds the whole mesh and particle set):
//  * mesh tables keep the reference's element/edge numbering
//    (cylag/geom/triangular_mesh.pxd:3-45) so that walk order, boundary edges
//    and tie-breaking are the reference's;
//  * Eulerian fields are uploaded once for every time record; the step works
//    on time-interpolated node arrays (eulerian_field_set.pyx:866-938);
//  * particles are structure-of-arrays in a physical order that is re-sorted
//    by element; `pool` holds the reference's pool index of each slot, which
//    fixes the particle's place in the serial random stream.
#pragma once
#include <stdint.h>
#include "cyl_rng.h"

namespace cyl {

constexpr double EPS = 1.e-16;             // constants.pyx:4
constexpr double EPS_LSM2 = 1.e-8;         // constants.pyx:7
constexpr double EPS_REBOUND = 1.e-8;      // constants.pyx:18
constexpr double DEG_TO_RAD = 0.017453292519943295;
constexpr double GRAV = 9.80665;
constexpr double NU0 = 1.3e-6;
constexpr double C0KOLM = 1.2;
constexpr double CMU = 0.09;
constexpr double KAPPA = 0.41;
constexpr int PART_LOC_O = -1;
constexpr int PART_LOC_B = -2;
constexpr int PART_LOC_A = -3;
constexpr int MAX_OPT = 8;

// error codes reported by kernels (first failing pool index is kept)
enum ErrCode : int {
  E_OK = 0,
  E_NOT_IN_TRI = 1,       // "Failed interpolation (.), point not in triangulation"
  E_BELOW_BOTTOM = 2,
  E_ABOVE_SURFACE = 3,
  E_DEGENERATE_LAYER = 4,
  E_LSM2_COEF = 5,        // "LSM2 requires finite a1 >= 0, dt >= 0 and a1*dt"
  E_SCHEME = 6,
  E_RNG_WINDOW = 7,       // random-stream window too small (host retries)
  E_MISPREDICT = 8,       // draw count differs from the uniform prediction (host reruns)
  E_METHOD1 = 9,          // neighbourhood search requested (reference has no table for it)
  E_LSM3 = 10,
  E_RNG_EVENTS = 11,      // a slow Ziggurat start missing from the step's event table (internal)
  E_LSM3_COEF = 12,       // "LSM3 requires finite a1 >= 0, A2, TL > 0 and dt >= 0"
  E_LSM3_CHOL = 13,       // "LSM3 covariance has a nonzero singular row / negative pivot"
};

struct MeshDev {
  int nv, nt, ne, nopen;
  const double* x;
  const double* y;
  const int* tri;          // nt*3
  const int* edges;        // ne*2
  const int* edgetri;      // ne*2
  const int* triedge;      // nt*3
  const int* edge_labels;  // ne
  const double* det;       // nt*3
  const double* gx;        // nt*3
  const double* gy;        // nt*3
  const double* centroid;  // nt*2
  // uniform bucket grid over triangle bounding boxes (fallback localization)
  double g_x0, g_y0, g_inv;
  int g_nx, g_ny;
  const int* g_start;      // g_nx*g_ny+1
  const int* g_items;
};

struct FieldsDev {
  int dim, nlayers, npoints, nopt;
  int fid_zs, fid_ux, fid_uy, fid_uz, fid_turbeng, fid_dissip, fid_zb;
  int frozen, prev_valid;
  int friction_model;
  double friction_coefficient;
  const double* tmp[4 + MAX_OPT];  // time-interpolated node arrays by field id
  const double* prev[4];           // previous interpolation of zs, u, v, w
};

// Particle pool, structure of arrays (one buffer of the double-buffered state).
struct PoolDev {
  double* x;
  double* y;
  double* z;
  double* up[3];   // velocity
  double* us[3];   // fluid velocity seen
  int* tri;
  int* ll;
  int* inactive;
};

// Per-slot data that the step does not modify.
struct SlotDev {
  const int* pool;      // reference pool index of the slot
  const int* rank;      // uniform mode: index of the slot's first draw unit / units-per-particle
  const int64_t* off;   // general mode: first draw unit of the slot (exclusive scan), or null
  const int* scen;      // scenario of the slot (batched runs) or null
  const double* diameter;
  const double* volume;
};

struct ParticleConst {
  double rho0, density, drag_coef, added_mass_coef, buoyancy_velocity_value;
  int drag_coef_model, added_mass_force, buoyancy_velocity_model;
  double a0c, a1c, a2c, a3c;
};

struct ParamsDev {
  int model, time_scheme, particle_velocity_init, rng_method;
  int diffusion_model, diffusion_lsm2_option;
  double horizontal_diffusivity, vertical_diffusivity, schmidt_number;
  int boundary_conditions, boundary_conditions_type, max_wall_rebounds, bnd_statistics;
  double rebound_damping_coef, wall_roughness_angle, stranding_probability;
  int compute_depth;
  int lsm3_edge_opt;
  double lsm3_tl_horizontal, lsm3_tl_vertical;
  int stranding_output;
  int bottom_deposition;  // deposition on the bed instead of the bottom reflection (extension)
};

// Random stream of one scenario for one step.
struct StreamDev {
  uint64_t base;        // state before the step's first draw
  uint64_t inc;
  int units_per_part;   // uniform mode: draw units per active particle (0 = general mode)
  // Ziggurat / polar: compacted slow events in the step window
  const int64_t* ev_key;    // normal index (zig) or key r_j - j (polar) of each event
  const int64_t* ev_extra;  // cumulative extra positions (zig)
  const double* ev_val;     // value of the slow normal of each event (zig)
  int64_t n_ev;
  int64_t window;           // positions covered by the event scan
  // bucket index over ev_key: ev_bucket[b] = #events with key < (b << ev_bshift)
  const int* ev_bucket;
  int64_t n_bucket;
  int ev_bshift;
};

struct StepDev {
  double dt, time;
  int nstream;
  const StreamDev* streams;   // per scenario
  const JumpTable* jumps;     // per scenario
  int* bnd_stats;             // [nscen][nopen] for this step or null
  int* n_deleted;             // counter
  unsigned long long* err;    // (pool << 8 | code) min-reduced, per scenario
  int* mispredict;            // flag
  int* units_used;            // per particle draw units (general count pass) or null
};

}  // namespace cyl
