// Engine interface shared by the core driver and the refinement-node code.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/qz.h"

namespace qz
{

constexpr double kPi = 3.14159265358979323846; // global_variables.c:221

void set_error(const char *fmt, ...);
bool hip_ok(hipError_t e, const char *what);
extern thread_local std::string g_last_error;

struct Timer
{
  struct Rec
  {
    std::string name;
    hipEvent_t a, b;
    double bytes;
  };
  struct Acc
  {
    std::string name;
    double ms;
    int count;
    double bytes; // algorithmic bytes (each input read once, each output written once)
  };
  bool on = false;      // per-kernel events (kernels launched eagerly)
  bool host_on = false; // host wall-time spans only (graph replays kept)
  std::vector<hipEvent_t> pending; // open begin() events; timed regions may nest
  std::vector<Rec> recs;
  std::vector<Acc> acc;
  std::mutex host_mu; // host spans may be recorded by the set-up thread too
  void begin(hipStream_t s);
  void end(hipStream_t s, const char *name, double bytes = 0.0);
  void collect();
  void host(const char *name, double ms); // host wall time of a stage part
};

// Host wall time of a scope, recorded as "host:<name>" while timing is on.
struct HostSpan
{
  Timer &t;
  const char *name;
  std::chrono::steady_clock::time_point t0;
  HostSpan(Timer &t_, const char *n) : t(t_), name(n), t0(std::chrono::steady_clock::now()) {}
  ~HostSpan()
  {
    if (t.on || t.host_on)
      t.host(name, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  }
};

struct TreeHost; // host tree (qz_tree.h)
struct TreeDev;  // device mirror of the tree
struct BmgPlan;  // node multigrid set-up of one tree level
struct TreeSnap; // a recorded tree state (adjoint records)

struct TreeState
{
  int n_nodes = 0;
  TreeHost *host = nullptr;
  TreeDev *dev = nullptr;
};

struct EngineBase
{
  qz_config cfg{};
  hipStream_t stream{};
  Timer timer;
  qz_stats stats{};
  int N = 0;       // particles
  int n = 0;       // head cells per side
  int D = 0;       // octree key depth below the head
  int end_bit = 0; // sorted key bits
  double total_mass = 0.0;
  size_t bytes = 0;
  TreeState *tree = nullptr;

  virtual ~EngineBase() {}
  virtual bool init() = 0;
  virtual bool set_particles(const double *const *src, bool init_head) = 0;
  virtual bool get_particles(double *const *dst) = 0;
  virtual bool update_A(double dt) = 0;
  virtual bool update_B(double dt) = 0;
  virtual bool reset() = 0;
  virtual bool tree_adaptation() = 0;
  virtual bool grid_density() = 0;
  virtual bool potential_head() = 0;
  virtual bool potential_branches() = 0;
  virtual bool grid_acceleration() = 0;
  virtual bool particle_acceleration() = 0;
  virtual bool head_fields(double *out5) = 0;
  virtual bool set_head_potential(const double *pot) = 0;
  virtual bool observables(int exact_pe, double *out6) = 0;
  virtual bool timestep(double cfl, double dt_max, double *dt) = 0;
  virtual bool force_vjp(const double *abar, double *xbar) = 0;
  virtual bool adjoint_forward(int nsteps, double dt, int every) = 0;
  virtual bool adjoint_backward(const double *xbN, const double *vbN, double *xb0, double *vb0) = 0;
  virtual int num_nodes() = 0;
  virtual bool node_header(int i, int *hdr) = 0;
  virtual bool node_lists(int i, int *ixyz, int *bxyz, int *sxyz) = 0;
  virtual bool node_fields(int i, double *out5) = 0;
  virtual bool bench_sweep(int reps, double *ms, double *bytes) = 0;
  virtual bool write_snapshot(const char *path, const qz_units &u, double actualtime, bool with_obs) = 0;
  virtual bool write_snapshot_async(const char *path, const qz_units &u, double actualtime, const double *obs6) = 0;
  virtual bool snapshot_wait() = 0;
  virtual bool copy_probe(size_t bytes, double *ms) = 0;
  virtual bool join_ahead() = 0; // wait for host work run ahead on a thread
  virtual bool force_overlapped(bool with_particles) = 0; // qz_force with the tree built during the head solve
};

EngineBase *make_engine(const qz_config &cfg);

} // namespace qz
