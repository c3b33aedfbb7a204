// PlotFile.cpp -- WritePlotFile and RecoverParticles (Source/IO.cpp:11-63) in
// the AMReX plotfile layout.
//
// The grid is a single-level plotfile of the valid cells:
//   <dir>/Header                plotfile header (HyperCLaw-V1.1)
//   <dir>/Level_0/Cell_H        FabArray header: box, FAB file, per-component min/max
//   <dir>/Level_0/Cell_D_00000  "FAB (...)" header line, then the doubles of
//                               the box, component-major, x fastest
// The particles are an AMReX particle checkpoint (is_checkpoint = 1), the
// layout the reference's readers parse (Scripts/visualization/
// amrex_plot_tools.py:89-210, Scripts/tests/EmuReader.py):
//   <dir>/neutrinos/Header               version, dim, real names, int names,
//                                        checkpoint flag, count, next id, levels,
//                                        grids per level, (file, count, offset)
//   <dir>/neutrinos/Level_0/DATA_00000   count x (id, cpu) int32, then
//                                        count x (pos_x, pos_y, pos_z, PIdx...) double
//   <dir>/neutrinos/Level_0/Particle_H   the particle grids' box array
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <limits>
#include <numeric>
#include <sstream>

#include "Emu.hpp"

namespace emu {

namespace {

std::vector<std::string> herm_names(int nf, const std::string& pre, const std::string& post)
{
    std::vector<std::string> out;
    for (int i = 0; i < nf; i++)
        for (int j = i; j < nf; j++) {
            out.push_back(pre + std::to_string(i) + std::to_string(j) + "_Re" + post);
            if (j > i) out.push_back(pre + std::to_string(i) + std::to_string(j) + "_Im" + post);
        }
    return out;
}

std::string box_string(const int n[3])
{
    std::ostringstream s;
    s << "((0,0,0) (" << n[0] - 1 << "," << n[1] - 1 << "," << n[2] - 1 << ") (0,0,0))";
    return s.str();
}

void make_dir(const std::string& d)
{
    std::error_code ec;
    std::filesystem::create_directories(d, ec);
    if (ec) Abort("cannot create directory '" + d + "': " + ec.message());
}

std::ofstream open_out(const std::string& path, bool binary = false)
{
    std::ofstream f(path, binary ? std::ios::binary : std::ios::out);
    if (!f) Abort("cannot write '" + path + "'");
    return f;
}

} // namespace

std::vector<std::string> grid_component_names(int nf)
{
    // GIdx order (Evolve.H:18-22 with its generated fill)
    std::vector<std::string> n = {"rho", "T", "Ye"};
    for (const char* v : {"N", "Fx", "Fy", "Fz"})
        for (const char* t : {"", "bar"}) {
            auto h = herm_names(nf, v, t);
            n.insert(n.end(), h.begin(), h.end());
        }
    return n;
}

std::vector<std::string> particle_attribute_names(int nf)
{
    // PIdx order (FlavoredNeutrinoContainer.H:12-26 with its generated fill)
    std::vector<std::string> n = {"time", "x", "y", "z", "pupx", "pupy", "pupz", "pupt"};
    for (const char* t : {"", "bar"}) {
        n.push_back(std::string("N") + t);
        n.push_back(std::string("L") + t);
        auto h = herm_names(nf, "f", t);
        n.insert(n.end(), h.begin(), h.end());
    }
    return n;
}

void write_plotfile_grid(const std::string& dir, const PlotFileSpec& sp, const double* valid)
{
    const auto names = grid_component_names(sp.nf);
    const int ncomp = (int)names.size();
    const int64_t ncells = (int64_t)sp.ncell[0] * sp.ncell[1] * sp.ncell[2];
    make_dir(dir + "/Level_0");
    {
        auto h = open_out(dir + "/Header");
        h << std::setprecision(17);
        h << "HyperCLaw-V1.1\n" << ncomp << "\n";
        for (const auto& n : names) h << n << "\n";
        h << 3 << "\n" << sp.time << "\n" << 0 << "\n";
        h << sp.plo[0] << " " << sp.plo[1] << " " << sp.plo[2] << "\n";
        h << sp.phi[0] << " " << sp.phi[1] << " " << sp.phi[2] << "\n";
        h << "\n"; // refinement ratios: none for a single level
        h << box_string(sp.ncell) << " \n";
        h << sp.step << " \n";
        for (int d = 0; d < 3; d++) h << (sp.phi[d] - sp.plo[d]) / sp.ncell[d] << " ";
        h << "\n" << 0 << "\n" << 0 << "\n"; // cartesian coordinates, boundary width 0
        h << 0 << " " << 1 << " " << sp.time << "\n" << sp.step << "\n";
        for (int d = 0; d < 3; d++) h << sp.plo[d] << " " << sp.phi[d] << "\n";
        h << "Level_0/Cell\n";
    }
    {
        auto f = open_out(dir + "/Level_0/Cell_D_00000", true);
        std::ostringstream hd;
        // IEEE double, little-endian byte order
        hd << "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))" << box_string(sp.ncell) << " " << ncomp
           << "\n";
        const std::string s = hd.str();
        f.write(s.data(), (std::streamsize)s.size());
        f.write((const char*)valid, (std::streamsize)(ncells * ncomp * sizeof(double)));
        if (!f) Abort("short write to '" + dir + "/Level_0/Cell_D_00000'");
    }
    {
        auto c = open_out(dir + "/Level_0/Cell_H");
        c << std::setprecision(17);
        c << 1 << "\n" << 0 << "\n" << ncomp << "\n" << 0 << "\n";
        c << "(1 0\n" << box_string(sp.ncell) << "\n)\n";
        c << 1 << "\n" << "FabOnDisk: Cell_D_00000 0\n\n";
        std::vector<double> mn(ncomp), mx(ncomp);
        for (int k = 0; k < ncomp; k++) {
            const double* v = valid + (size_t)k * ncells;
            double a = std::numeric_limits<double>::infinity(), b = -a;
            for (int64_t i = 0; i < ncells; i++) {
                a = std::min(a, v[i]);
                b = std::max(b, v[i]);
            }
            mn[k] = a;
            mx[k] = b;
        }
        c << 1 << "," << ncomp << "\n";
        for (double v : mn) c << v << ",";
        c << "\n" << 1 << "," << ncomp << "\n";
        for (double v : mx) c << v << ",";
        c << "\n";
    }
}

void write_plotfile_particles(const std::string& dir, const std::string& name, const PlotFileSpec& sp, int64_t np,
                              const int64_t* ids, const double* pos, const double* rdata)
{
    const auto names = particle_attribute_names(sp.nf);
    const int na = (int)names.size();
    const std::string base = dir + "/" + name;
    make_dir(base + "/Level_0");
    // particles in id order, as one grid covering the domain
    std::vector<int64_t> order((size_t)np);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return ids[a] < ids[b]; });
    int64_t maxid = 0;
    for (int64_t i = 0; i < np; i++) maxid = std::max(maxid, ids[i]);
    {
        auto h = open_out(base + "/Header");
        h << "Version_Two_Dot_Zero_double\n" << 3 << "\n" << na << "\n";
        for (const auto& n : names) h << n << "\n";
        h << 0 << "\n" << 1 << "\n" << np << "\n" << maxid + 1 << "\n" << 0 << "\n";
        h << 1 << "\n" << 0 << " " << np << " " << 0 << "\n";
    }
    {
        auto p = open_out(base + "/Level_0/Particle_H");
        p << "(1 0\n" << box_string(sp.ncell) << "\n)\n";
    }
    {
        auto f = open_out(base + "/Level_0/DATA_00000", true);
        std::vector<int32_t> iv((size_t)np * 2);
        for (int64_t i = 0; i < np; i++) {
            iv[2 * i] = (int32_t)ids[order[i]];
            iv[2 * i + 1] = 0;
        }
        f.write((const char*)iv.data(), (std::streamsize)(iv.size() * sizeof(int32_t)));
        std::vector<double> rv((size_t)np * (3 + na));
        for (int64_t i = 0; i < np; i++) {
            double* r = rv.data() + (size_t)i * (3 + na);
            for (int d = 0; d < 3; d++) r[d] = pos[(size_t)d * np + order[i]];
            for (int k = 0; k < na; k++) r[3 + k] = rdata[(size_t)k * np + order[i]];
        }
        f.write((const char*)rv.data(), (std::streamsize)(rv.size() * sizeof(double)));
        if (!f) Abort("short write to '" + base + "/Level_0/DATA_00000'");
    }
}

void read_plotfile_header(const std::string& dir, double& time, int& step)
{
    std::ifstream h(dir + "/Header");
    if (!h) Abort("cannot read '" + dir + "/Header'");
    std::string line;
    std::getline(h, line); // version
    int ncomp = 0;
    h >> ncomp;
    std::getline(h, line);
    for (int k = 0; k < ncomp; k++) std::getline(h, line);
    int dim = 0, finest = 0;
    h >> dim >> time >> finest;
    double v;
    for (int d = 0; d < 2 * dim; d++) h >> v; // prob_lo, prob_hi
    std::getline(h, line);
    std::getline(h, line); // refinement ratios
    std::getline(h, line); // domain box
    h >> step;
    if (!h) Abort("malformed plotfile header in '" + dir + "'");
}

int64_t read_plotfile_particles(const std::string& dir, const std::string& name, int nf, std::vector<int64_t>& ids,
                                std::vector<double>& pos, std::vector<double>& rdata)
{
    const std::string base = dir + "/" + name;
    std::ifstream h(base + "/Header");
    if (!h) Abort("cannot read '" + base + "/Header'");
    std::string version;
    std::getline(h, version);
    if (version.size() < 7 || version.substr(version.size() - 7) != "_double")
        Abort("particle data in '" + base + "' are not double precision");
    int dim = 0, nreal = 0, nint = 0, checkpoint = 0, finest = 0;
    int64_t np = 0, nextid = 0;
    h >> dim >> nreal;
    std::string s;
    std::getline(h, s);
    std::vector<std::string> names((size_t)nreal);
    for (auto& n : names) std::getline(h, n);
    h >> nint;
    std::getline(h, s);
    for (int k = 0; k < nint; k++) std::getline(h, s);
    h >> checkpoint >> np >> nextid >> finest;
    if (!h || dim != 3 || finest != 0) Abort("unsupported particle header in '" + base + "'");
    if (names != particle_attribute_names(nf))
        Abort("particle attributes in '" + base + "' do not match NUM_FLAVORS = " + std::to_string(nf));
    const int ni = checkpoint ? 2 + nint : 0, nr = 3 + nreal;
    int ngrids = 0;
    h >> ngrids;
    ids.assign((size_t)np, 0);
    pos.assign((size_t)3 * np, 0.0);
    rdata.assign((size_t)nreal * np, 0.0);
    int64_t ip = 0;
    for (int gi = 0; gi < ngrids; gi++) {
        int which = 0;
        int64_t count = 0, where = 0;
        h >> which >> count >> where;
        if (!h) Abort("malformed particle header in '" + base + "'");
        if (count == 0) continue;
        if (ip + count > np) Abort("particle counts in '" + base + "' exceed the total");
        char fn[64];
        std::snprintf(fn, sizeof fn, "/Level_0/DATA_%05d", which);
        std::ifstream f(base + fn, std::ios::binary);
        if (!f) Abort("cannot read '" + base + fn + "'");
        f.seekg(where);
        std::vector<int32_t> iv((size_t)count * ni);
        std::vector<double> rv((size_t)count * nr);
        f.read((char*)iv.data(), (std::streamsize)(iv.size() * sizeof(int32_t)));
        f.read((char*)rv.data(), (std::streamsize)(rv.size() * sizeof(double)));
        if (!f) Abort("short read of '" + base + fn + "'");
        for (int64_t i = 0; i < count; i++) {
            ids[ip + i] = ni ? iv[(size_t)i * ni] : ip + i + 1;
            for (int d = 0; d < 3; d++) pos[(size_t)d * np + ip + i] = rv[(size_t)i * nr + d];
            for (int k = 0; k < nreal; k++) rdata[(size_t)k * np + ip + i] = rv[(size_t)i * nr + 3 + k];
        }
        ip += count;
    }
    if (ip != np) Abort("particle file '" + base + "' holds fewer particles than its header says");
    return np;
}

// ---------------------------------------------------------------------------
// IO.cpp:11-36 and 38-63
// ---------------------------------------------------------------------------
void WritePlotFile(const MultiFab& state, const FlavoredNeutrinoContainer& neutrinos, const Geometry& geom,
                   double time, int step, int write_plot_particles)
{
    char dirname[32];
    std::snprintf(dirname, sizeof dirname, "plt%05d", step); // amrex::Concatenate("plt", step)
    const std::string dir = dirname;
    const TestParams& p = geom.parms();
    PlotFileSpec sp{};
    sp.nf = p.nf;
    for (int d = 0; d < 3; d++) {
        sp.ncell[d] = p.ncell[d];
        sp.plo[d] = 0.0;
    }
    sp.phi[0] = p.Lx;
    sp.phi[1] = p.Ly;
    sp.phi[2] = p.Lz;
    sp.time = time;
    sp.step = step;
    // the valid cells of the fab
    const std::vector<double> fab = state.to_host();
    const emu_grid* g = state.fab();
    const int nx = g->hi[0] - g->lo[0] + 1, ny = g->hi[1] - g->lo[1] + 1, nz = g->hi[2] - g->lo[2] + 1;
    const int ncomp = g->ncomp;
    const int64_t ncells = (int64_t)p.ncell[0] * p.ncell[1] * p.ncell[2];
    std::vector<double> valid((size_t)ncomp * ncells);
    for (int c = 0; c < ncomp; c++)
        for (int k = 0; k < p.ncell[2]; k++)
            for (int j = 0; j < p.ncell[1]; j++)
                for (int i = 0; i < p.ncell[0]; i++)
                    valid[(((size_t)c * p.ncell[2] + k) * p.ncell[1] + j) * p.ncell[0] + i] =
                        fab[(((size_t)c * nz + (k - g->lo[2])) * ny + (j - g->lo[1])) * nx + (i - g->lo[0])];
    write_plotfile_grid(dir, sp, valid.data());
    if (write_plot_particles == 1) {
        std::vector<double> pos, rdata;
        std::vector<int64_t> ids;
        neutrinos.to_host(pos, rdata, ids);
        write_plotfile_particles(dir, "neutrinos", sp, neutrinos.TotalNumberOfParticles(), ids.data(), pos.data(),
                                 rdata.data());
    }
    auto j = open_out(dir + "/job_info");
    j << "Emu job information (MI355X build)\n";
    j << "geometry.prob_lo: 0 0 0\ngeometry.prob_hi: " << p.Lx << " " << p.Ly << " " << p.Lz << "\n";
    j << "amr.n_cell: " << p.ncell[0] << " " << p.ncell[1] << " " << p.ncell[2] << "\n";
}

void RecoverParticles(const std::string& dir, FlavoredNeutrinoContainer& neutrinos, double& time, int& step)
{
    read_plotfile_header(dir, time, step);
    std::vector<double> pos, rdata;
    std::vector<int64_t> ids;
    const int64_t np = read_plotfile_particles(dir, "neutrinos", neutrinos.geom().nf(), ids, pos, rdata);
    neutrinos.from_host(np, pos, rdata, ids);
}

} // namespace emu
