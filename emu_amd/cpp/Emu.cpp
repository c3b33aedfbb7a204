// Emu.cpp -- the C++ call surface of Emu.hpp over libemu_amd.so.
#include "Emu.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <numeric>
#include <sstream>

namespace emu {

namespace {

bool g_throw = false;

// CGSUnitsConst::eV and PhysConst::c (Source/Constants.H:6-26)
constexpr double kEV = 1.60218e-12;
constexpr double kC = 2.99792458e10;
constexpr double kC2 = kC * kC;

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) Abort(std::string(what) + ": " + hipGetErrorString(e), EMU_ERR_HIP);
}

void emu_check(int code, const char* what, emu_ctx* ctx, int64_t bad = -1)
{
    if (code == EMU_OK) return;
    std::string msg = std::string(what) + " failed (status " + std::to_string(code) + ")";
    if (ctx) {
        const char* le = emu_last_error(ctx);
        if (le && *le) msg += ": " + std::string(le);
    }
    if (bad >= 0) msg += " at particle " + std::to_string(bad);
    Abort(msg, code, bad);
}

std::string trim(const std::string& s)
{
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

// ParmParse-style table: key -> raw value text
struct Table {
    std::map<std::string, std::string> kv;

    bool has(const std::string& k) const { return kv.count(k) != 0; }
    const std::string& raw(const std::string& k) const
    {
        auto it = kv.find(k);
        // ParmParse::get aborts on a missing key
        if (it == kv.end()) Abort("ParmParse::get: required key '" + k + "' not found in the inputs");
        return it->second;
    }
    double real(const std::string& k) const
    {
        const std::string& v = raw(k);
        char* e = nullptr;
        const double x = std::strtod(v.c_str(), &e);
        if (e == v.c_str()) Abort("ParmParse: key '" + k + "' expects a number, got '" + v + "'");
        return x;
    }
    int integer(const std::string& k) const
    {
        const double x = real(k);
        if (x != std::floor(x)) Abort("ParmParse: key '" + k + "' expects an integer, got '" + raw(k) + "'");
        return (int)x;
    }
    bool boolean(const std::string& k) const
    {
        const std::string v = raw(k);
        if (v == "true" || v == "1") return true;
        if (v == "false" || v == "0") return false;
        Abort("ParmParse: key '" + k + "' expects a boolean, got '" + v + "'");
    }
    std::string str(const std::string& k) const
    {
        std::string v = raw(k);
        if (v.size() >= 2 && v.front() == '"' && v.back() == '"') v = v.substr(1, v.size() - 2);
        return v;
    }
    void ivect(const std::string& k, int out[3]) const
    {
        std::string v = raw(k);
        for (char& ch : v)
            if (ch == '(' || ch == ')' || ch == ',') ch = ' ';
        std::istringstream is(v);
        for (int d = 0; d < 3; d++)
            if (!(is >> out[d])) Abort("ParmParse: key '" + k + "' expects three integers, got '" + raw(k) + "'");
    }
    void add_line(const std::string& line0)
    {
        std::string line = line0.substr(0, line0.find('#'));
        const size_t eq = line.find('=');
        if (eq == std::string::npos) return;
        const std::string k = trim(line.substr(0, eq));
        if (!k.empty()) kv[k] = trim(line.substr(eq + 1));
    }
};

} // namespace

void set_throw_on_error(bool enable) { g_throw = enable; }

void Abort(const std::string& msg, int code, int64_t bad_index)
{
    if (g_throw) throw Error(msg, code, bad_index);
    std::fprintf(stderr, "emu::Abort: %s\n", msg.c_str());
    std::fflush(stderr);
    std::abort();
}

// ---------------------------------------------------------------------------
// TestParams (Parameters.H:53-130): the required keys and the conversions in
// the reference's order
// ---------------------------------------------------------------------------
void TestParams::Initialize(const std::string& inputs_file, const std::vector<std::string>& overrides)
{
    Table t;
    std::ifstream in(inputs_file);
    if (!in) Abort("cannot open inputs file '" + inputs_file + "'");
    std::string line;
    while (std::getline(in, line)) t.add_line(line);
    for (const auto& o : overrides) t.add_line(o); // command-line key=value, as AMReX's ParmParse
    Initialize(t.kv);
}

void TestParams::Initialize(const std::map<std::string, std::string>& table)
{
    Table t;
    t.kv = table;
    // compile-time settings of the reference (Make.Emu), run-time keys here
    if (t.has("nf")) nf = t.integer("nf");
    if (t.has("shape_order")) shape_order = t.integer("shape_order");
    if (nf != 2 && nf != 3) Abort("nf must be 2 or 3", EMU_ERR_UNSUPPORTED);
    if (shape_order < 0 || shape_order > 2) Abort("shape_order must be 0, 1 or 2", EMU_ERR_UNSUPPORTED);
    if (t.has("directions")) directions = t.str("directions");
    if (t.has("ndirections")) ndirections = t.integer("ndirections");
    if (t.has("seed")) seed = (uint64_t)std::strtoull(t.raw("seed").c_str(), nullptr, 10);
    if (t.has("integration.type")) integration_type = t.integer("integration.type");
    if (t.has("integration.rk.type")) integration_rk_type = t.integer("integration.rk.type");
    (void)Tableau::preset(integration_type, integration_rk_type); // refuses unknown types

    simulation_type = t.integer("simulation_type");
    t.ivect("ncell", ncell);
    Lx = t.real("Lx");
    Ly = t.real("Ly");
    Lz = t.real("Lz");
    t.ivect("nppc", nppc);
    nphi_equator = t.integer("nphi_equator");
    max_grid_size = t.integer("max_grid_size");
    nsteps = t.integer("nsteps");
    end_time = t.real("end_time");
    rho_in = t.real("rho_g_ccm");
    Ye_in = t.real("Ye");
    T_in = t.real("T_MeV");
    cfl_factor = t.real("cfl_factor");
    flavor_cfl_factor = t.real("flavor_cfl_factor");
    max_adaptive_speedup = t.real("max_adaptive_speedup");
    write_plot_every = t.integer("write_plot_every");
    write_plot_particles_every = t.integer("write_plot_particles_every");
    do_restart = t.boolean("do_restart");
    restart_dir = t.str("restart_dir");
    maxError = t.real("maxError");

    mass1 = t.real("mass1_eV");
    mass2 = t.real("mass2_eV");
    theta12 = t.real("theta12_degrees");
    alpha1 = t.real("alpha1_degrees");
    mass1 *= kEV / kC2;
    mass2 *= kEV / kC2;
    theta12 *= M_PI / 180.;
    alpha1 *= M_PI / 180.;
    // the reference reads these for every NUM_FLAVORS >= 2
    mass3 = t.real("mass3_eV");
    theta13 = t.real("theta13_degrees");
    theta23 = t.real("theta23_degrees");
    alpha2 = t.real("alpha2_degrees");
    deltaCP = t.real("deltaCP_degrees");
    mass3 *= kEV / kC2;
    theta13 *= M_PI / 180.;
    theta23 *= M_PI / 180.;
    alpha2 *= M_PI / 180.;
    deltaCP *= M_PI / 180.;

    if (simulation_type == 3) {
        st3_amplitude = t.real("st3_amplitude");
        st3_wavelength_fraction_of_domain = t.integer("st3_wavelength_fraction_of_domain");
    }
    if (simulation_type == 4) {
        st4_theta = t.real("st4_theta");
        st4_thetabar = t.real("st4_thetabar");
        st4_phi = t.real("st4_phi");
        st4_phibar = t.real("st4_phibar");
        st4_ndens = t.real("st4_ndens");
        st4_ndensbar = t.real("st4_ndensbar");
        st4_fluxfac = t.real("st4_fluxfac");
        st4_fluxfacbar = t.real("st4_fluxfacbar");
        st4_amplitude = t.real("st4_amplitude");
    }
    if (simulation_type == 5) {
        st5_nnue = t.real("st5_nnue");
        st5_nnua = t.real("st5_nnua");
        st5_nnux = t.real("st5_nnux");
        st5_fxnue = t.real("st5_fxnue");
        st5_fxnua = t.real("st5_fxnua");
        st5_fxnux = t.real("st5_fxnux");
        st5_fynua = t.real("st5_fynua");
        st5_fynux = t.real("st5_fynux");
        st5_fynue = t.real("st5_fynue");
        st5_fznue = t.real("st5_fznue");
        st5_fznua = t.real("st5_fznua");
        st5_fznux = t.real("st5_fznux");
        st5_amplitude = t.real("st5_amplitude");
    }
}

emu_params TestParams::c_params() const
{
    emu_params p{};
    p.nf = nf;
    p.shape_order = shape_order;
    for (int d = 0; d < 3; d++) {
        p.ncell[d] = ncell[d];
        p.plo[d] = 0.0;
    }
    p.phi[0] = Lx;
    p.phi[1] = Ly;
    p.phi[2] = Lz;
    p.rho = rho_in;
    p.Ye = Ye_in;
    p.T = T_in;
    p.mass[0] = mass1;
    p.mass[1] = mass2;
    p.mass[2] = mass3;
    p.theta12 = theta12;
    p.theta13 = theta13;
    p.theta23 = theta23;
    p.alpha1 = alpha1;
    p.alpha2 = alpha2;
    p.deltaCP = deltaCP;
    p.maxError = maxError;
    p.cfl_factor = cfl_factor;
    p.flavor_cfl_factor = flavor_cfl_factor;
    p.max_adaptive_speedup = max_adaptive_speedup;
    return p;
}

emu_init_config TestParams::init_config() const
{
    emu_init_config c{};
    c.simulation_type = simulation_type;
    for (int d = 0; d < 3; d++) c.nppc[d] = nppc[d];
    if (directions == "sphere") c.direction_set = EMU_DIRS_SPHERE;
    else if (directions == "cube26") c.direction_set = EMU_DIRS_CUBE26;
    else if (directions == "fibonacci") c.direction_set = EMU_DIRS_FIBONACCI;
    else Abort("directions must be sphere, cube26 or fibonacci, got '" + directions + "'");
    c.nphi_equator = nphi_equator;
    c.ndirections = ndirections;
    c.st3_wavelength_fraction_of_domain = st3_wavelength_fraction_of_domain;
    c.st3_amplitude = st3_amplitude;
    c.st4_ndens = st4_ndens;
    c.st4_theta = st4_theta;
    c.st4_phi = st4_phi;
    c.st4_fluxfac = st4_fluxfac;
    c.st4_ndensbar = st4_ndensbar;
    c.st4_thetabar = st4_thetabar;
    c.st4_phibar = st4_phibar;
    c.st4_fluxfacbar = st4_fluxfacbar;
    c.st4_amplitude = st4_amplitude;
    const double st5[13] = {st5_nnue,  st5_nnua,  st5_nnux,  st5_fxnue, st5_fxnua, st5_fxnux,    st5_fynue,
                            st5_fynua, st5_fynux, st5_fznue, st5_fznua, st5_fznux, st5_amplitude};
    std::memcpy(c.st5, st5, sizeof st5);
    c.seed = seed;
    return c;
}

// ---------------------------------------------------------------------------
// Geometry
// ---------------------------------------------------------------------------
Geometry::Geometry(const TestParams& parms) : parms_(parms)
{
    const emu_params p = parms_.c_params();
    emu_check(emu_create(&p, &ctx_), "emu_create", nullptr);
}

Geometry::~Geometry()
{
    if (ctx_) emu_destroy(ctx_);
}

void Geometry::sync() const { emu_check(emu_sync(ctx_), "emu_sync", ctx_); }

// ---------------------------------------------------------------------------
// MultiFab
// ---------------------------------------------------------------------------
MultiFab::MultiFab(const Geometry& geom) : geom_(geom)
{
    int ng[3];
    emu_check(emu_grid_ngrow(geom.ctx(), ng), "emu_grid_ngrow", geom.ctx());
    const TestParams& p = geom.parms();
    cells_ = 1;
    for (int d = 0; d < 3; d++) {
        grid_.lo[d] = -ng[d];
        grid_.hi[d] = p.ncell[d] - 1 + ng[d];
        cells_ *= p.ncell[d] + 2 * ng[d];
    }
    grid_.ncomp = geom.ncomp();
    n_ = cells_ * grid_.ncomp;
    hip_check(hipMalloc((void**)&grid_.data, (size_t)n_ * sizeof(double)), "MultiFab alloc");
    // main.cpp:80-85
    setVal(0.0);
    setVal(p.rho_in, 0, 1);
    setVal(p.T_in, 1, 1);
    setVal(p.Ye_in, 2, 1);
    FillBoundary(geom);
}

MultiFab::~MultiFab()
{
    if (grid_.data) (void)hipFree(grid_.data);
}

void MultiFab::setVal(double v) { setVal(v, 0, grid_.ncomp); }

void MultiFab::setVal(double v, int comp, int ncomp)
{
    if (comp < 0 || ncomp < 0 || comp + ncomp > grid_.ncomp) Abort("MultiFab::setVal: component range");
    geom_.sync();
    std::vector<double> h((size_t)cells_ * ncomp, v);
    hip_check(hipMemcpy(grid_.data + (size_t)comp * cells_, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice),
              "MultiFab::setVal");
}

void MultiFab::FillBoundary(const Geometry& geom)
{
    emu_check(emu_fill_boundary(geom.ctx(), &grid_), "FillBoundary", geom.ctx());
}

std::vector<double> MultiFab::to_host() const
{
    geom_.sync();
    std::vector<double> h((size_t)n_);
    hip_check(hipMemcpy(h.data(), grid_.data, h.size() * sizeof(double), hipMemcpyDeviceToHost), "MultiFab::to_host");
    return h;
}

// ---------------------------------------------------------------------------
// FlavoredNeutrinoContainer
// ---------------------------------------------------------------------------
double FlavoredNeutrinoContainer::Vvac_max = 0.0;

FlavoredNeutrinoContainer::FlavoredNeutrinoContainer(const Geometry& geom) : geom_(geom) { resize(0); }

FlavoredNeutrinoContainer::~FlavoredNeutrinoContainer()
{
    if (block_) (void)hipFree(block_);
    if (ids_) (void)hipFree(ids_);
}

void FlavoredNeutrinoContainer::resize(int64_t np)
{
    if (np == np_ && (block_ || np == 0)) return;
    geom_.sync();
    if (block_) (void)hipFree(block_);
    if (ids_) (void)hipFree(ids_);
    block_ = nullptr;
    ids_ = nullptr;
    np_ = np;
    const int na = geom_.nattribs();
    std::memset(&view_, 0, sizeof view_);
    view_.np = np;
    if (np == 0) return;
    hip_check(hipMalloc((void**)&block_, (size_t)np * (3 + na) * sizeof(double)), "particle alloc");
    hip_check(hipMalloc((void**)&ids_, (size_t)np * sizeof(int64_t)), "particle id alloc");
    hip_check(hipMemset(block_, 0, (size_t)np * (3 + na) * sizeof(double)), "particle clear");
    for (int d = 0; d < 3; d++) view_.pos[d] = block_ + (size_t)d * np;
    for (int k = 0; k < na; k++) view_.rdata[k] = block_ + (size_t)(3 + k) * np;
}

void FlavoredNeutrinoContainer::InitParticles(const TestParams* parms)
{
    const emu_params p = parms->c_params();
    const emu_init_config cfg = parms->init_config();
    const int64_t n = emu_init_count(&p, &cfg);
    if (n < 0) Abort("InitParticles: invalid configuration", EMU_ERR_ARG);
    resize(n);
    emu_check(emu_init_particles(geom_.ctx(), &cfg, &view_, ids_), "InitParticles", geom_.ctx());
    // FlavoredNeutrinoContainerInit.cpp:657-660
    emu_check(emu_vvac_max(geom_.ctx(), &view_, &Vvac_max), "Vvac_max", geom_.ctx());
}

void FlavoredNeutrinoContainer::copyParticles(const FlavoredNeutrinoContainer& other, bool /*local*/)
{
    resize(other.np_);
    if (np_ == 0) return;
    emu_check(emu_copy_particles(geom_.ctx(), &view_, &other.view_), "copyParticles", geom_.ctx());
    geom_.sync();
    hip_check(hipMemcpy(ids_, other.ids_, (size_t)np_ * sizeof(int64_t), hipMemcpyDeviceToDevice), "copy ids");
}

void FlavoredNeutrinoContainer::rk_combine(const FlavoredNeutrinoContainer& src, int n, const double* a,
                                           const FlavoredNeutrinoContainer* const* F)
{
    resize(src.np_);
    if (np_ == 0) return;
    std::vector<emu_particles> views((size_t)std::max(n, 1));
    for (int j = 0; j < n; j++) views[j] = F[j]->view_;
    emu_check(emu_rk_combine(geom_.ctx(), &view_, &src.view_, n, a, views.data()), "rk_combine", geom_.ctx());
    // the ids travel with the copy, in stream order (the context's stream is the null stream)
    hip_check(hipMemcpyAsync(ids_, src.ids_, (size_t)np_ * sizeof(int64_t), hipMemcpyDeviceToDevice, nullptr),
              "copy ids");
}

void FlavoredNeutrinoContainer::SyncLocation(int type)
{
    if (type != Sync::CoordinateToPosition && type != Sync::PositionToCoordinate) Abort("SyncLocation: bad type");
    emu_check(emu_sync_location(geom_.ctx(), &view_, type), "SyncLocation", geom_.ctx());
}

void FlavoredNeutrinoContainer::RedistributeLocal()
{
    emu_check(emu_redistribute_local(geom_.ctx(), &view_), "RedistributeLocal", geom_.ctx());
}

void FlavoredNeutrinoContainer::Renormalize(const TestParams* /*parms*/)
{
    int64_t bad = -1;
    const int code = emu_renormalize(geom_.ctx(), &view_, &bad);
    if (code == EMU_ERR_TRACE) emu_check(code, "Renormalize: the trace of f is too far from 1", geom_.ctx(), bad);
    if (code == EMU_ERR_NEGDIAG) emu_check(code, "Renormalize: a diagonal of f is too negative", geom_.ctx(), bad);
    if (code == EMU_ERR_LENGTH) emu_check(code, "Renormalize: the length of f is too far from L", geom_.ctx(), bad);
    emu_check(code, "Renormalize", geom_.ctx(), bad);
}

void FlavoredNeutrinoContainer::apply_rhs(double a, const FlavoredNeutrinoContainer& dFdt)
{
    emu_check(emu_apply_rhs(geom_.ctx(), &view_, a, &dFdt.view_), "ApplyFlavoredNeutrinoRHS", geom_.ctx());
}

void FlavoredNeutrinoContainer::from_host(int64_t np, const std::vector<double>& pos, const std::vector<double>& rdata,
                                          const std::vector<int64_t>& ids)
{
    const int na = geom_.nattribs();
    if ((int64_t)pos.size() != 3 * np || (int64_t)rdata.size() != na * np || (int64_t)ids.size() != np)
        Abort("FlavoredNeutrinoContainer::from_host: array sizes");
    resize(np);
    if (np == 0) return;
    hip_check(hipMemcpy(block_, pos.data(), pos.size() * sizeof(double), hipMemcpyHostToDevice), "from_host pos");
    hip_check(hipMemcpy(block_ + (size_t)3 * np, rdata.data(), rdata.size() * sizeof(double), hipMemcpyHostToDevice),
              "from_host rdata");
    hip_check(hipMemcpy(ids_, ids.data(), ids.size() * sizeof(int64_t), hipMemcpyHostToDevice), "from_host ids");
}

void FlavoredNeutrinoContainer::to_host(std::vector<double>& pos, std::vector<double>& rdata,
                                        std::vector<int64_t>& ids) const
{
    geom_.sync();
    const int na = geom_.nattribs();
    pos.assign((size_t)3 * np_, 0.0);
    rdata.assign((size_t)na * np_, 0.0);
    ids.assign((size_t)np_, 0);
    if (np_ == 0) return;
    hip_check(hipMemcpy(pos.data(), block_, pos.size() * sizeof(double), hipMemcpyDeviceToHost), "to_host pos");
    hip_check(hipMemcpy(rdata.data(), block_ + (size_t)3 * np_, rdata.size() * sizeof(double), hipMemcpyDeviceToHost),
              "to_host rdata");
    hip_check(hipMemcpy(ids.data(), ids_, ids.size() * sizeof(int64_t), hipMemcpyDeviceToHost), "to_host ids");
}

// ---------------------------------------------------------------------------
// Evolve.H
// ---------------------------------------------------------------------------
void deposit_to_mesh(const FlavoredNeutrinoContainer& neutrinos, MultiFab& state, const Geometry& geom)
{
    emu_check(emu_deposit_to_mesh(geom.ctx(), neutrinos.view(), state.fab()), "deposit_to_mesh", geom.ctx());
}

void interpolate_rhs_from_mesh(FlavoredNeutrinoContainer& neutrinos_rhs, const MultiFab& state, const Geometry& geom,
                               const TestParams* /*parms*/)
{
    emu_check(emu_interpolate_rhs_from_mesh(geom.ctx(), neutrinos_rhs.view(), state.fab()),
              "interpolate_rhs_from_mesh", geom.ctx());
}

double compute_dt(const Geometry& geom, double cfl_factor, const MultiFab& state,
                  const FlavoredNeutrinoContainer& /*neutrinos*/, double flavor_cfl_factor,
                  double max_adaptive_speedup)
{
    // the context holds the factors it was created with; other values would
    // silently be ignored, so they are refused
    const TestParams& p = geom.parms();
    if (cfl_factor != p.cfl_factor || flavor_cfl_factor != p.flavor_cfl_factor ||
        max_adaptive_speedup != p.max_adaptive_speedup)
        Abort("compute_dt: CFL factors differ from the Geometry's parameters", EMU_ERR_UNSUPPORTED);
    double dt = 0.0;
    emu_check(emu_compute_dt(geom.ctx(), state.fab(), FlavoredNeutrinoContainer::Vvac_max, &dt), "compute_dt",
              geom.ctx());
    return dt;
}

// ---------------------------------------------------------------------------
// TimeIntegrator, RK4
// ---------------------------------------------------------------------------
Tableau Tableau::preset(int integration_type, int rk_type)
{
    Tableau t{};
    if (integration_type == 0) rk_type = 1; // the ForwardEuler integrator
    else if (integration_type != 1) Abort("integration.type must be 0 or 1", EMU_ERR_UNSUPPORTED);
    switch (rk_type) {
    case 1: // forward Euler
        t.nstages = 1;
        t.b[0] = 1.0;
        break;
    case 2: // trapezoid
        t.nstages = 2;
        t.A[1][0] = 1.0;
        t.b[0] = 0.5;
        t.b[1] = 0.5;
        t.c[1] = 1.0;
        break;
    case 3: // SSPRK3
        t.nstages = 3;
        t.A[1][0] = 1.0;
        t.A[2][0] = 0.25;
        t.A[2][1] = 0.25;
        t.b[0] = 1. / 6.;
        t.b[1] = 1. / 6.;
        t.b[2] = 2. / 3.;
        t.c[1] = 1.0;
        t.c[2] = 0.5;
        break;
    case 4: // classical RK4
        t.nstages = 4;
        t.A[1][0] = 0.5;
        t.A[2][1] = 0.5;
        t.A[3][2] = 1.0;
        t.b[0] = 1. / 6.;
        t.b[1] = 1. / 3.;
        t.b[2] = 1. / 3.;
        t.b[3] = 1. / 6.;
        t.c[1] = 0.5;
        t.c[2] = 0.5;
        t.c[3] = 1.0;
        break;
    default: Abort("integration.rk.type must be 1, 2, 3 or 4", EMU_ERR_UNSUPPORTED);
    }
    return t;
}

bool Tableau::is_rk4() const
{
    const Tableau r = preset(1, 4);
    return std::memcmp(this, &r, sizeof r) == 0;
}

TimeIntegrator::TimeIntegrator(FlavoredNeutrinoContainer& old_data, FlavoredNeutrinoContainer& new_data, double time,
                               int step, const Tableau& tableau)
    : tab_(tableau), old_(&old_data), new_(&new_data), time_(time), step_(step)
{
    for (int i = 0; i < tab_.nstages; i++) F_.push_back(new FlavoredNeutrinoContainer(old_data.geom()));
}

TimeIntegrator::~TimeIntegrator()
{
    for (auto* f : F_) delete f;
}

void TimeIntegrator::advance(double dt)
{
    // each stage state and the result: copyParticles(old) followed by one
    // ApplyFlavoredNeutrinoRHS per earlier stage, as one pass (rk_combine)
    double a[4];
    for (int i = 0; i < tab_.nstages; i++) {
        for (int j = 0; j < i; j++) a[j] = dt * tab_.A[i][j];
        new_->rk_combine(*old_, i, a, F_.data());
        rhs_(*F_[i], *new_, time_ + tab_.c[i] * dt);
    }
    for (int i = 0; i < tab_.nstages; i++) a[i] = dt * tab_.b[i];
    new_->rk_combine(*old_, tab_.nstages, a, F_.data());
    time_ += dt;
}

void TimeIntegrator::integrate(double start_dt, double end_time, int nsteps)
{
    // as AMReX's loop: the post-timestep hook sees the number of the step just
    // taken counted from the start step (0 for the first step of a new run)
    dt_ = start_dt;
    for (int n = 0; n < nsteps && time_ < end_time; n++) {
        advance(dt_);
        if (post_) post_();
        step_ += 1;
        std::swap(old_, new_);
    }
}

// ---------------------------------------------------------------------------
// output
// ---------------------------------------------------------------------------
void write_dump(const std::string& path, const FlavoredNeutrinoContainer& c, double time)
{
    std::vector<double> pos, rdata;
    std::vector<int64_t> ids;
    c.to_host(pos, rdata, ids);
    const int64_t np = c.TotalNumberOfParticles();
    const int na = c.geom().nattribs();
    std::vector<int64_t> order((size_t)np);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return ids[a] < ids[b]; });
    std::ofstream out(path, std::ios::binary);
    if (!out) Abort("cannot write '" + path + "'");
    const int32_t nf = c.geom().nf(), na32 = na;
    out.write("EMUDUMP1", 8);
    out.write((const char*)&nf, 4);
    out.write((const char*)&na32, 4);
    out.write((const char*)&np, 8);
    out.write((const char*)&time, 8);
    std::vector<int64_t> oid((size_t)np);
    for (int64_t i = 0; i < np; i++) oid[i] = ids[order[i]];
    out.write((const char*)oid.data(), np * 8);
    std::vector<double> col((size_t)np);
    for (int d = 0; d < 3; d++) {
        for (int64_t i = 0; i < np; i++) col[i] = pos[(size_t)d * np + order[i]];
        out.write((const char*)col.data(), np * 8);
    }
    for (int k = 0; k < na; k++) {
        for (int64_t i = 0; i < np; i++) col[i] = rdata[(size_t)k * np + order[i]];
        out.write((const char*)col.data(), np * 8);
    }
    if (!out) Abort("short write to '" + path + "'");
}

// ---------------------------------------------------------------------------
// evolve_flavor (main.cpp:37-227)
// ---------------------------------------------------------------------------
EvolveResult evolve_flavor(const TestParams* parms, const EvolveOptions& opt)
{
    Geometry geom(*parms);
    MultiFab state(geom);
    FlavoredNeutrinoContainer neutrinos_old(geom);
    FlavoredNeutrinoContainer neutrinos_new(geom);
    EvolveResult res;
    if (!opt.quiet) std::cout << "Initializing particles... ";
    double initial_time = 0.0;
    int initial_step = 0;
    if (parms->do_restart) {
        RecoverParticles(parms->restart_dir, neutrinos_old, initial_time, initial_step);
        // the reference sets Vvac_max only in InitParticles, so a restarted run
        // (a fresh process there) uses its zero initial value
        FlavoredNeutrinoContainer::Vvac_max = 0.0;
        if (!opt.quiet)
            std::cout << "Restarting after time step: " << initial_step - 1 << " t = " << initial_time
                      << " s.  ct = " << kC * initial_time << " cm" << std::endl;
    } else {
        neutrinos_old.InitParticles(parms);
    }
    neutrinos_new.copyParticles(neutrinos_old, true);
    deposit_to_mesh(neutrinos_old, state, geom);
    if (opt.write_plotfiles && !parms->do_restart) WritePlotFile(state, neutrinos_old, geom, initial_time, initial_step, 1);
    if (!opt.quiet) std::cout << "Done. " << std::endl;
    res.np = neutrinos_old.TotalNumberOfParticles();
    double run_fom = 0.0;
    // main.cpp:181-191 (a write_plot_every of 0 writes no grid-only plotfiles)
    auto plot_due = [&](int next, int& with_particles) {
        with_particles = parms->write_plot_particles_every > 0 && next % parms->write_plot_particles_every == 0;
        return opt.write_plotfiles &&
               ((parms->write_plot_every > 0 && next % parms->write_plot_every == 0) || with_particles);
    };

    if (opt.path == "plain") {
        TimeIntegrator integrator(neutrinos_old, neutrinos_new, initial_time, initial_step,
                                  Tableau::preset(parms->integration_type, parms->integration_rk_type));
        auto source_fun = [&](FlavoredNeutrinoContainer& neutrinos_rhs, const FlavoredNeutrinoContainer& neutrinos,
                              double /*time*/) {
            deposit_to_mesh(neutrinos, state, geom);
            state.FillBoundary(geom);
            neutrinos_rhs.copyParticles(neutrinos, true);
            interpolate_rhs_from_mesh(neutrinos_rhs, state, geom, parms);
        };
        auto post_timestep_fun = [&]() {
            auto& neutrinos = integrator.get_new_data();
            neutrinos.SyncLocation(Sync::CoordinateToPosition);
            neutrinos.RedistributeLocal();
            neutrinos.SyncLocation(Sync::PositionToCoordinate);
            neutrinos.Renormalize(parms);
            const int step = integrator.get_step_number();
            const double time = integrator.get_time();
            if (!opt.quiet)
                std::cout << "Completed time step: " << step << " t = " << time << " s.  ct = " << kC * time << " cm"
                          << std::endl;
            run_fom += (double)neutrinos.TotalNumberOfParticles();
            int wpp = 0;
            if (plot_due(step + 1, wpp)) WritePlotFile(state, neutrinos, geom, time, step + 1, wpp);
            const double dt = compute_dt(geom, parms->cfl_factor, state, neutrinos, parms->flavor_cfl_factor,
                                         parms->max_adaptive_speedup);
            integrator.set_timestep(dt);
            res.dt = dt;
        };
        integrator.set_rhs(source_fun);
        integrator.set_post_timestep(post_timestep_fun);
        const double starting_dt = compute_dt(geom, parms->cfl_factor, state, neutrinos_old,
                                              parms->flavor_cfl_factor, parms->max_adaptive_speedup);
        if (!opt.quiet) std::cout << "Starting timestepping loop... " << std::endl;
        geom.sync();
        const auto t0 = std::chrono::steady_clock::now();
        integrator.integrate(starting_dt, parms->end_time, parms->nsteps);
        geom.sync();
        res.run_time_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        res.steps = integrator.get_step_number() - initial_step;
        res.time = integrator.get_time();
        if (!opt.dump_path.empty()) write_dump(opt.dump_path, integrator.get_old_data(), res.time);
    } else if (opt.path == "fused-fp64" || opt.path == "fused-fp32") {
        if (!Tableau::preset(parms->integration_type, parms->integration_rk_type).is_rk4())
            Abort("the fused step is classical RK4 (integration.type = 1, integration.rk.type = 4)",
                  EMU_ERR_UNSUPPORTED);
        const int precision = opt.path == "fused-fp64" ? EMU_PLAIN_FP64 : EMU_FP32_BULK;
        emu_state* st = nullptr;
        emu_check(emu_state_create(geom.ctx(), res.np, precision, &st), "emu_state_create", geom.ctx());
        struct Closer {
            emu_state* s;
            ~Closer() { emu_state_destroy(s); }
        } closer{st};
        emu_check(emu_state_load(st, neutrinos_old.view(), neutrinos_old.ids()), "emu_state_load", geom.ctx());
        if (parms->do_restart)
            emu_check(emu_state_set_vvac_max(st, FlavoredNeutrinoContainer::Vvac_max), "emu_state_set_vvac_max",
                      geom.ctx());
        if (!opt.quiet) std::cout << "Starting timestepping loop... " << std::endl;
        double time = 0, dt = 0;
        int64_t steps = 0, bad = -1;
        geom.sync();
        const auto t0 = std::chrono::steady_clock::now();
        time = initial_time;
        for (int n = 0; n < parms->nsteps && time < parms->end_time; n++) {
            emu_check(emu_state_evolve(st, 1, 0.0), "emu_state_evolve", geom.ctx());
            const int code = emu_state_info(st, &time, &dt, &steps, &bad);
            if (code != EMU_OK)
                emu_check(code,
                          code == EMU_ERR_MOVE ? "fused step: a particle moved more than one cell"
                          : code == EMU_ERR_DT ? "compute_dt"
                                               : "Renormalize",
                          geom.ctx(), bad);
            run_fom += (double)res.np;
            const int step = initial_step + (int)steps - 1;
            if (!opt.quiet)
                std::cout << "Completed time step: " << step << " t = " << time << " s.  ct = " << kC * time << " cm"
                          << std::endl;
            int wpp = 0;
            if (plot_due(step + 1, wpp)) {
                // the fused state's grid holds combined fields: deposit the step's
                // result in the reference's layout for the plotfile
                emu_check(emu_state_store(st, neutrinos_new.view(), neutrinos_new.ids()), "emu_state_store",
                          geom.ctx());
                deposit_to_mesh(neutrinos_new, state, geom);
                WritePlotFile(state, neutrinos_new, geom, time, step + 1, wpp);
            }
        }
        geom.sync();
        res.run_time_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        res.steps = (int)steps;
        res.time = time;
        res.dt = dt;
        if (!opt.dump_path.empty()) {
            emu_check(emu_state_store(st, neutrinos_old.view(), neutrinos_old.ids()), "emu_state_store", geom.ctx());
            write_dump(opt.dump_path, neutrinos_old, res.time);
        }
    } else {
        Abort("unknown path '" + opt.path + "' (plain, fused-fp64, fused-fp32)");
    }
    res.fom = res.run_time_s > 0 ? run_fom / res.run_time_s / 1.e6 : 0.0;
    if (!opt.quiet) {
        std::cout << "Done. " << std::endl;
        std::printf("Run time w/o initialization (seconds) = %.3f\n", res.run_time_s);
        std::printf("Average number of particles advanced per microsecond = %.3f\n", res.fom);
    }
    return res;
}

} // namespace emu
