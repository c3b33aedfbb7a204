// Emu.hpp -- the reference's C++ call surface for the flavor-evolution step,
// over the C ABI of libemu_amd.so.
//
// Names, argument meaning and error behaviour follow the reference:
//   TestParams                 Source/Parameters.H:11-131
//   FlavoredNeutrinoContainer  Source/FlavoredNeutrinoContainer.H:49-92
//   deposit_to_mesh, interpolate_rhs_from_mesh, compute_dt
//                              Source/Evolve.H:29-33
//   TimeIntegrator (RK4)       the AMReX TimeIntegrator that main.cpp:129 builds
//   evolve_flavor              Source/main.cpp:37-227
// so a caller written against the reference's step compiles against this
// header with the AMReX types replaced by the ones below.  Particle and grid
// data live in device memory (SoA, PIdx / GIdx order); every computation runs
// in the HIP kernels behind include/emu_amd.h.
//
// Errors: where the reference calls amrex::Error / amrex::Abort (which end the
// process), emu::Abort prints the message and calls std::abort(), unless
// set_throw_on_error(true) was called, in which case it throws emu::Error
// carrying the library's status code and the offending particle index.
#pragma once

#include <cstdint>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "emu_amd.h"
#include "emu_init.h"

namespace emu {

// Sync::CoordinateToPosition / PositionToCoordinate (FlavoredNeutrinoContainer.H:28-31)
struct Sync {
    enum { CoordinateToPosition = 0, PositionToCoordinate = 1 };
};

struct Error : std::runtime_error {
    int code;
    int64_t bad_index;
    Error(const std::string& msg, int code_, int64_t bad) : std::runtime_error(msg), code(code_), bad_index(bad) {}
};

void set_throw_on_error(bool enable);
[[noreturn]] void Abort(const std::string& msg, int code = EMU_ERR_ARG, int64_t bad_index = -1);

// ---------------------------------------------------------------------------
// Parameters.H:11-131.  Initialize() reads an AMReX ParmParse deck ("key =
// value", '#' comments, tuples as "(a,b,c)") plus "key=value" overrides, with
// the reference's required keys and its unit conversions (masses eV -> g,
// degrees -> radians).  NUM_FLAVORS and SHAPE_FACTOR_ORDER, compile-time in
// the reference (Make.Emu:1-2), are the optional keys "nf" and
// "shape_order"; "directions" (sphere | cube26 | fibonacci), "ndirections"
// and "seed" select the direction set and the random stream of the initial
// conditions.
// ---------------------------------------------------------------------------
struct TestParams {
    int nf = 2, shape_order = 2;
    int ncell[3] = {1, 1, 1};
    int nppc[3] = {1, 1, 1};
    int nphi_equator = 1;
    double Lx = 1, Ly = 1, Lz = 1;
    int max_grid_size = 64;
    int nsteps = 1;
    double end_time = 1.0;
    int write_plot_every = 0, write_plot_particles_every = 0;
    double rho_in = 0, Ye_in = 1, T_in = 10;
    int simulation_type = 0;
    double cfl_factor = 0.5, flavor_cfl_factor = 0.5, max_adaptive_speedup = 0;
    bool do_restart = false;
    std::string restart_dir;
    double maxError = 1e-6;
    double mass1 = 0, mass2 = 0, mass3 = 0;          // grams
    double theta12 = 0, theta13 = 0, theta23 = 0;    // radians
    double alpha1 = 0, alpha2 = 0, deltaCP = 0;      // radians
    int st3_wavelength_fraction_of_domain = 1;
    double st3_amplitude = 0;
    double st4_ndens = 0, st4_theta = 0, st4_phi = 0, st4_fluxfac = 0;
    double st4_ndensbar = 0, st4_thetabar = 0, st4_phibar = 0, st4_fluxfacbar = 0;
    double st4_amplitude = 0;
    double st5_nnue = 0, st5_nnua = 0, st5_nnux = 0;
    double st5_fxnue = 0, st5_fxnua = 0, st5_fxnux = 0;
    double st5_fynue = 0, st5_fynua = 0, st5_fynux = 0;
    double st5_fznue = 0, st5_fznua = 0, st5_fznux = 0;
    double st5_amplitude = 0;
    std::string directions = "sphere";
    int ndirections = 0;
    uint64_t seed = 1234;
    // the AMReX TimeIntegrator's ParmParse keys: integration.type (0 forward
    // Euler, 1 Runge-Kutta) and integration.rk.type (1 Euler, 2 trapezoid,
    // 3 SSPRK3, 4 RK4)
    int integration_type = 1;
    int integration_rk_type = 4;

    void Initialize(const std::string& inputs_file, const std::vector<std::string>& overrides = {});
    void Initialize(const std::map<std::string, std::string>& table);
    emu_params c_params() const;
    emu_init_config init_config() const;
};

// ---------------------------------------------------------------------------
// amrex::Geometry of the single periodic box, together with the library
// context it needs (mass matrix, stream, scratch).
// ---------------------------------------------------------------------------
class Geometry {
public:
    explicit Geometry(const TestParams& parms);
    ~Geometry();
    Geometry(const Geometry&) = delete;
    Geometry& operator=(const Geometry&) = delete;

    emu_ctx* ctx() const { return ctx_; }
    const TestParams& parms() const { return parms_; }
    int nf() const { return parms_.nf; }
    int nattribs() const { return 12 + 2 * parms_.nf * parms_.nf; }
    int ncomp() const { return 3 + 8 * parms_.nf * parms_.nf; }
    void sync() const;

private:
    TestParams parms_;
    emu_ctx* ctx_ = nullptr;
};

// ---------------------------------------------------------------------------
// The MultiFab `state` (main.cpp:78): GIdx::ncomp components, ngrow ghost
// cells, one fab covering the domain (device memory, Array4 order).
// ---------------------------------------------------------------------------
class MultiFab {
public:
    explicit MultiFab(const Geometry& geom);
    ~MultiFab();
    MultiFab(const MultiFab&) = delete;
    MultiFab& operator=(const MultiFab&) = delete;

    void setVal(double v);                          // every component
    void setVal(double v, int comp, int ncomp);     // components comp .. comp+ncomp-1
    void FillBoundary(const Geometry& geom);        // periodic ghost fill (main.cpp:85,137)
    emu_grid* fab() { return &grid_; }
    const emu_grid* fab() const { return &grid_; }
    int64_t size() const { return n_; }             // doubles in the fab
    std::vector<double> to_host() const;

private:
    const Geometry& geom_;
    emu_grid grid_{};
    int64_t n_ = 0;
    int64_t cells_ = 0;
};

// ---------------------------------------------------------------------------
// FlavoredNeutrinoContainer (FlavoredNeutrinoContainer.H:49-92): particles
// as device SoA arrays, pos[3] and rdata[PIdx::nattribs], with particle ids.
// ---------------------------------------------------------------------------
class FlavoredNeutrinoContainer {
public:
    explicit FlavoredNeutrinoContainer(const Geometry& geom);
    ~FlavoredNeutrinoContainer();
    FlavoredNeutrinoContainer(const FlavoredNeutrinoContainer&) = delete;
    FlavoredNeutrinoContainer& operator=(const FlavoredNeutrinoContainer&) = delete;

    void InitParticles(const TestParams* parms);                           // FlavoredNeutrinoContainerInit.cpp:136-661
    void copyParticles(const FlavoredNeutrinoContainer& other, bool local); // main.cpp:115,148
    void SyncLocation(int type);                                          // FlavoredNeutrinoContainer.cpp:8-41
    void RedistributeLocal();                                             // FlavoredNeutrinoContainer.H:83-90
    void Renormalize(const TestParams* parms);                            // FlavoredNeutrinoContainer.cpp:89-113
    int64_t TotalNumberOfParticles() const { return np_; }

    // Vvac_max of the particles, set by InitParticles
    // (FlavoredNeutrinoContainerInit.cpp:657-660).
    static double Vvac_max;

    emu_particles* view() { return &view_; }
    const emu_particles* view() const { return &view_; }
    int64_t* ids() const { return ids_; }
    const Geometry& geom() const { return geom_; }

    // ApplyFlavoredNeutrinoRHS over the container (FlavoredNeutrinoContainer.H:33-47)
    void apply_rhs(double a, const FlavoredNeutrinoContainer& dFdt);
    // copyParticles(src) followed by apply_rhs(a[j], *F[j]) for j < n, in one pass
    void rk_combine(const FlavoredNeutrinoContainer& src, int n, const double* a,
                    const FlavoredNeutrinoContainer* const* F);
    // host copies (SoA, particle order), for output and restart
    void to_host(std::vector<double>& pos, std::vector<double>& rdata, std::vector<int64_t>& ids) const;
    void from_host(int64_t np, const std::vector<double>& pos, const std::vector<double>& rdata,
                   const std::vector<int64_t>& ids);
    void resize(int64_t np);

private:
    const Geometry& geom_;
    int64_t np_ = 0;
    double* block_ = nullptr;
    int64_t* ids_ = nullptr;
    emu_particles view_{};
};

// Evolve.H:29-33
void deposit_to_mesh(const FlavoredNeutrinoContainer& neutrinos, MultiFab& state, const Geometry& geom);
void interpolate_rhs_from_mesh(FlavoredNeutrinoContainer& neutrinos_rhs, const MultiFab& state, const Geometry& geom,
                               const TestParams* parms);
double compute_dt(const Geometry& geom, double cfl_factor, const MultiFab& state,
                  const FlavoredNeutrinoContainer& neutrinos, double flavor_cfl_factor, double max_adaptive_speedup);

// ---------------------------------------------------------------------------
// The AMReX TimeIntegrator with an explicit preset tableau (classical RK4,
// integration.type = 1 / integration.rk.type = 4, in every deck): stage i
// copies the old data, adds dt a_ij F_j for j < i, evaluates F_i; the new data
// is old + dt sum b_i F_i; then the post-timestep hook runs and old/new swap
// roles.  integration.type = 0 (ForwardEuler) is the one-stage tableau.
// ---------------------------------------------------------------------------
struct Tableau {
    int nstages;
    double A[4][4];
    double b[4];
    double c[4];
    static Tableau preset(int integration_type, int rk_type);
    bool is_rk4() const;
};

class TimeIntegrator {
public:
    using Rhs = std::function<void(FlavoredNeutrinoContainer& rhs, const FlavoredNeutrinoContainer& state, double time)>;
    TimeIntegrator(FlavoredNeutrinoContainer& old_data, FlavoredNeutrinoContainer& new_data, double time, int step,
                   const Tableau& tableau = Tableau::preset(1, 4));
    ~TimeIntegrator();
    void set_rhs(Rhs f) { rhs_ = std::move(f); }
    void set_post_timestep(std::function<void()> f) { post_ = std::move(f); }
    void set_timestep(double dt) { dt_ = dt; }
    void integrate(double start_dt, double end_time, int nsteps);
    FlavoredNeutrinoContainer& get_new_data() { return *new_; }
    FlavoredNeutrinoContainer& get_old_data() { return *old_; }
    int get_step_number() const { return step_; }
    double get_time() const { return time_; }

private:
    void advance(double dt);
    Tableau tab_;
    FlavoredNeutrinoContainer* old_;
    FlavoredNeutrinoContainer* new_;
    std::vector<FlavoredNeutrinoContainer*> F_;
    Rhs rhs_;
    std::function<void()> post_;
    double time_, dt_ = 0;
    int step_;
};

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// IO.cpp: plotfiles in the AMReX layout (the valid cells of `state` as a
// single-level plotfile, the particles as an AMReX particle checkpoint named
// "neutrinos") and the restart from one.  PlotFile.cpp holds the format.
// ---------------------------------------------------------------------------
struct PlotFileSpec {
    int nf;
    int ncell[3];
    double plo[3], phi[3];
    double time;
    int step;
};
std::vector<std::string> grid_component_names(int nf);      // GIdx::names
std::vector<std::string> particle_attribute_names(int nf);  // PIdx names
void write_plotfile_grid(const std::string& dir, const PlotFileSpec& sp, const double* valid);
void write_plotfile_particles(const std::string& dir, const std::string& name, const PlotFileSpec& sp, int64_t np,
                              const int64_t* ids, const double* pos, const double* rdata);
void read_plotfile_header(const std::string& dir, double& time, int& step);
int64_t read_plotfile_particles(const std::string& dir, const std::string& name, int nf, std::vector<int64_t>& ids,
                                std::vector<double>& pos, std::vector<double>& rdata);
void WritePlotFile(const MultiFab& state, const FlavoredNeutrinoContainer& neutrinos, const Geometry& geom,
                   double time, int step, int write_plot_particles); // IO.cpp:11-36
void RecoverParticles(const std::string& dir, FlavoredNeutrinoContainer& neutrinos, double& time,
                      int& step);                                   // IO.cpp:38-63

// evolve_flavor (main.cpp:37-227).  path = "plain": the reference's loop on
// the per-function calls above; "fused-fp64" / "fused-fp32": the same loop as
// whole RK4 steps of the cell-sorted state (emu_state_*).  Plotfiles follow
// write_plot_every / write_plot_particles_every as main.cpp:181-191 (the
// fused path deposits the step's result for their grid); do_restart reads
// the particles of restart_dir.  dump_path, if set, receives the final
// particles.
// ---------------------------------------------------------------------------
struct EvolveOptions {
    std::string path = "plain";
    bool quiet = false;
    std::string dump_path;
    bool write_plotfiles = true; // plt<step> directories in the working directory, as main.cpp
};

struct EvolveResult {
    int steps = 0;
    double time = 0, dt = 0;
    double run_time_s = 0;     // integrate() wall time, without initialization
    double fom = 0;            // particles advanced per microsecond (main.cpp:211-225)
    int64_t np = 0;
};

EvolveResult evolve_flavor(const TestParams* parms, const EvolveOptions& opt = {});

// Binary dump of a container, particles ordered by id:
//   "EMUDUMP1", int32 nf, int32 nattribs, int64 np, double time,
//   int64 ids[np], double pos[3][np], double rdata[nattribs][np].
void write_dump(const std::string& path, const FlavoredNeutrinoContainer& c, double time);

} // namespace emu
