// sim_tcpc_main.cpp - the sim_tcpc executable over liblbm355.
//
// Same command line, defaults, output tree and exit codes as
// solver/sim_tcpc/sim_tcpc.cu:25-125,788-857 driven through
// solver/lbm3d/core.h:1476-1628 (execute): reads the wall-mask geometry VTK,
// converts the material to lattice units, runs until the physical final time
// and writes results_<case_id>/log_rank000 and
// results_<case_id>/vtk3D/rank000_data_<n>.vtk.
// Exit codes: 0 finished, 1 usage / setup error, 2 terminate flag (zero
// inlet or outlet area).
//
// Extensions (new flags only): --collision, --precision, --steps,
// --inflow-profile, --ivc-umax/--svc-umax, --pulse-amp/--pulse-period,
// --cy-model, --stats-every and the materials cy-a, cy-b, blood-newtonian.
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/lbm355.h"

namespace {

struct Options {
    std::string geometry = "geometry.vtk";
    std::string material = "newtonian";
    std::string case_id = "tcpc_manual";
    double ivc_mls = 8.0;
    double svc_mls = 4.833333333;
    double final_time = 10.0;
    double vtk_period = 0.1;
    double print_period = 0.1;
    double stat_reset_period = -1.0;
    double lbm_viscosity = 1.0e-4;
    int block_size = 32;
    // extensions
    std::string collision = "cumulant";
    std::string precision = "fp32";
    std::string profile = "plug";
    std::string cy_model = "neq";
    int64_t steps = -1;
    double ivc_umax = -1.0, svc_umax = -1.0;
    double pulse_amp = 0.0, pulse_period = 0.0;
    int stats_every = 1;
    double max_wall_time = -1.0;   // the reference's wallTime (state.h:256), seconds
    bool save_at_end = false;
    bool autotune = true;
};

struct Material {
    std::string name;
    double density, nu_inf, mu0, lambda, a, n;
    bool carreau_yasuda;
};

// sim_tcpc.cu:52-68, plus the blood-like benchmark materials
Material material_by_name(const std::string& name)
{
    if (name == "newtonian") return {"newtonian", 1000.0, 1.0e-6, 1.0e-3, 1.0, 2.0, 1.0, false};
    if (name == "sx")
        return {"sx", 1151.0, 3.72863999716223e-3 / 1151.0, 3.32933884442747e-2, 1.20785676226298,
                0.925354152818718, 0.463126647, true};
    if (name == "gx")
        return {"gx", 1088.0, 4.85563692580026e-3 / 1088.0, 1.58517801201338e-2, 0.0280506051570583,
                2.34147778848562, 0.463126647, true};
    if (name == "blood-newtonian") return {"blood-newtonian", 1060.0, 0.0035 / 1060.0, 0.0035, 1.0, 2.0, 1.0, false};
    if (name == "cy-a") return {"cy-a", 1060.0, 0.00345 / 1060.0, 0.056, 3.313, 2.0, 0.3568, true};
    if (name == "cy-b") return {"cy-b", 1060.0, 0.0035 / 1060.0, 0.16, 8.2, 0.64, 0.2128, true};
    throw std::runtime_error("Unknown material '" + name + "'. Use newtonian, sx, or gx.");
}

void usage()
{
    std::cerr << "Usage: ./sim_tcpc --geometry geometry.vtk --material newtonian|sx|gx "
              << "--ivc-mls VALUE --svc-mls VALUE --case-id NAME [options]\n"
              << "Options:\n"
              << "  --final-time VALUE          Physical final time in seconds, default 10\n"
              << "  --vtk-period VALUE          3D VTK period in seconds, default 0.1\n"
              << "  --print-period VALUE        Log period in seconds, default 0.1\n"
              << "  --stat-reset-period VALUE   Statistics reset period, default disabled\n"
              << "  --lbm-viscosity VALUE       LBM reference viscosity, default 1e-4\n"
              << "  --block-size VALUE          accepted for compatibility (the brick layout has no y-block)\n"
              << "Extensions:\n"
              << "  --material cy-a|cy-b|blood-newtonian   benchmark materials\n"
              << "  --collision cumulant|bgk|mrt|kbc       default cumulant\n"
              << "  --precision fp32|fp64|fp16             population storage, default fp32 (fp16: f - w\n"
              << "                                         in IEEE half, stochastically rounded; admitted\n"
              << "                                         at --lbm-viscosity 1e-2 and above)\n"
              << "  --cy-model neq|off|fd-force            Carreau-Yasuda coupling, default neq; off = what a\n"
              << "                                         single-rank reference run computes, fd-force = the\n"
              << "                                         reference's multi-rank algorithm\n"
              << "  --steps N                              run exactly N steps (overrides --final-time)\n"
              << "  --inflow-profile plug|parabolic        default plug (u = flux / area)\n"
              << "  --ivc-umax V --svc-umax V              parabolic peak velocities in m/s\n"
              << "                                         (default: twice the plug mean velocity)\n"
              << "  --pulse-amp A --pulse-period T         u(t) = u (1 + A sin(2 pi t / T)), T in s\n"
              << "  --stats-every K                        statistics sum every K steps, default 1\n"
              << "  --max-wall-time S                      save results_<id>/current_state and stop after S\n"
              << "                                         seconds (creates the loadstate flag file)\n"
              << "  --save-state-at-end                    save results_<id>/current_state when the run ends\n"
              << "  --no-autotune                          keep the default launch configuration (by default it\n"
              << "                                         is timed on a copy of the case before the run)\n"
              << "Restart: a file results_<id>/loadstate loads results_<id>/current_state at start;\n"
              << "a file results_<id>/savestate saves it at the next print event.\n";
}

Options parse_args(int argc, char** argv)
{
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string arg(argv[i]);
        auto value = [&](const char* name) -> const char* {
            if (i + 1 >= argc) throw std::runtime_error(std::string("Missing value for ") + name);
            return argv[++i];
        };
        if (arg == "--help" || arg == "-h") { usage(); std::exit(0); }
        else if (arg == "--geometry") o.geometry = value("--geometry");
        else if (arg == "--material") o.material = value("--material");
        else if (arg == "--case-id" || arg == "--output-id") o.case_id = value(arg.c_str());
        else if (arg == "--ivc-mls") o.ivc_mls = std::atof(value("--ivc-mls"));
        else if (arg == "--svc-mls") o.svc_mls = std::atof(value("--svc-mls"));
        else if (arg == "--final-time") o.final_time = std::atof(value("--final-time"));
        else if (arg == "--vtk-period") o.vtk_period = std::atof(value("--vtk-period"));
        else if (arg == "--print-period") o.print_period = std::atof(value("--print-period"));
        else if (arg == "--stat-reset-period") o.stat_reset_period = std::atof(value("--stat-reset-period"));
        else if (arg == "--lbm-viscosity") o.lbm_viscosity = std::atof(value("--lbm-viscosity"));
        else if (arg == "--block-size") o.block_size = std::atoi(value("--block-size"));
        else if (arg == "--collision") o.collision = value("--collision");
        else if (arg == "--precision") o.precision = value("--precision");
        else if (arg == "--cy-model") o.cy_model = value("--cy-model");
        else if (arg == "--steps") o.steps = std::atoll(value("--steps"));
        else if (arg == "--inflow-profile") o.profile = value("--inflow-profile");
        else if (arg == "--ivc-umax") o.ivc_umax = std::atof(value("--ivc-umax"));
        else if (arg == "--svc-umax") o.svc_umax = std::atof(value("--svc-umax"));
        else if (arg == "--pulse-amp") o.pulse_amp = std::atof(value("--pulse-amp"));
        else if (arg == "--pulse-period") o.pulse_period = std::atof(value("--pulse-period"));
        else if (arg == "--stats-every") o.stats_every = std::atoi(value("--stats-every"));
        else if (arg == "--max-wall-time") o.max_wall_time = std::atof(value("--max-wall-time"));
        else if (arg == "--save-state-at-end") o.save_at_end = true;
        else if (arg == "--no-autotune") o.autotune = false;
        else throw std::runtime_error("Unknown argument '" + arg + "'");
    }
    return o;
}

// ---------------------------------------------------------------------------
// legacy VTK RECTILINEAR_GRID reader for the wall mask (the role of
// vtkRectilinearGridReader in readGeometry, sim_tcpc.cu:134-185): point
// dimensions - 1 = cells, dl = first positive x spacing, CELL_DATA "wall".
struct GeometryData {
    std::vector<int32_t> wall;
    int dim[3] = {0, 0, 0};
    double dx = 0.0;
};

class VtkText {
public:
    explicit VtkText(std::vector<char> d) : data_(std::move(d)) {}
    bool eof() const { return pos_ >= data_.size(); }
    std::string line()
    {
        size_t e = pos_;
        while (e < data_.size() && data_[e] != '\n') ++e;
        std::string s(data_.data() + pos_, e - pos_);
        pos_ = e < data_.size() ? e + 1 : e;
        while (!s.empty() && (s.back() == '\r' || s.back() == ' ' || s.back() == '\t')) s.pop_back();
        size_t b = 0;
        while (b < s.size() && (s[b] == ' ' || s[b] == '\t')) ++b;
        return s.substr(b);
    }
    std::string next_nonempty()
    {
        while (!eof()) {
            std::string s = line();
            if (!s.empty()) return s;
        }
        throw std::runtime_error("unexpected end of VTK file");
    }
    // n values of a VTK type, converted to double
    std::vector<double> values(size_t n, const std::string& type, bool binary)
    {
        std::vector<double> out(n);
        if (!binary) {
            size_t k = 0;
            while (k < n) {
                std::string s = next_nonempty();
                const char* p = s.c_str();
                char* end = nullptr;
                while (k < n) {
                    double v = std::strtod(p, &end);
                    if (end == p) break;
                    out[k++] = v;
                    p = end;
                }
            }
            return out;
        }
        size_t w = type_size(type);
        if (pos_ + n * w > data_.size()) throw std::runtime_error("truncated binary VTK array");
        const unsigned char* p = reinterpret_cast<const unsigned char*>(data_.data() + pos_);
        for (size_t i = 0; i < n; ++i) out[i] = decode(p + i * w, type);
        pos_ += n * w;
        if (pos_ < data_.size() && data_[pos_] == '\n') ++pos_;
        return out;
    }

private:
    static size_t type_size(const std::string& t)
    {
        if (t == "double" || t == "long" || t == "unsigned_long" || t == "vtktypeint64") return 8;
        if (t == "float" || t == "int" || t == "unsigned_int") return 4;
        if (t == "short" || t == "unsigned_short") return 2;
        if (t == "char" || t == "unsigned_char" || t == "bit") return 1;
        throw std::runtime_error("unsupported VTK data type " + t);
    }
    static double decode(const unsigned char* p, const std::string& t)
    {
        uint64_t v = 0;
        const size_t w = type_size(t);
        for (size_t i = 0; i < w; ++i) v = (v << 8) | p[i];  // big-endian
        if (t == "double") { double d; std::memcpy(&d, &v, 8); return d; }
        if (t == "float") { uint32_t u = uint32_t(v); float f; std::memcpy(&f, &u, 4); return f; }
        if (t == "int") return double(int32_t(uint32_t(v)));
        if (t == "unsigned_int") return double(uint32_t(v));
        if (t == "short") return double(int16_t(uint16_t(v)));
        if (t == "unsigned_short") return double(uint16_t(v));
        if (t == "char") return double(int8_t(uint8_t(v)));
        if (t == "long" || t == "vtktypeint64") return double(int64_t(v));
        return double(v);
    }
    std::vector<char> data_;
    size_t pos_ = 0;
};

std::vector<std::string> split(const std::string& s)
{
    std::vector<std::string> t;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) ++i;
        size_t j = i;
        while (j < s.size() && s[j] != ' ' && s[j] != '\t') ++j;
        if (j > i) t.push_back(s.substr(i, j - i));
        i = j;
    }
    return t;
}

std::string upper(std::string s)
{
    for (char& c : s) c = char(std::toupper(static_cast<unsigned char>(c)));
    return s;
}

GeometryData read_geometry(const std::string& file_name)
{
    std::ifstream in(file_name, std::ios::binary);
    if (!in) throw std::runtime_error(file_name + " is not a legacy VTK RectilinearGrid file.");
    std::vector<char> bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    VtkText r(std::move(bytes));
    r.line();  // version
    r.line();  // title
    const bool binary = upper(r.line()).rfind("BINARY", 0) == 0;
    std::vector<std::string> ds = split(r.next_nonempty());
    if (ds.size() < 2 || upper(ds[1]) != "RECTILINEAR_GRID")
        throw std::runtime_error(file_name + " is not a legacy VTK RectilinearGrid file.");
    GeometryData g;
    int pdim[3] = {0, 0, 0};
    std::vector<double> xc;
    size_t ncells = 0, npoints = 0;
    bool in_cells = false, have_wall = false;
    while (!r.eof()) {
        std::string s;
        try { s = r.next_nonempty(); } catch (const std::exception&) { break; }
        std::vector<std::string> t = split(s);
        const std::string key = upper(t[0]);
        if (key == "DIMENSIONS" && t.size() >= 4) {
            for (int i = 0; i < 3; ++i) pdim[i] = std::atoi(t[i + 1].c_str());
        } else if (key.size() == 13 && key.substr(1) == "_COORDINATES" && t.size() >= 3) {
            std::vector<double> c = r.values(size_t(std::atoll(t[1].c_str())), t[2], binary);
            if (key[0] == 'X') xc = c;
        } else if (key == "CELL_DATA" && t.size() >= 2) {
            ncells = size_t(std::atoll(t[1].c_str()));
            in_cells = true;
        } else if (key == "POINT_DATA" && t.size() >= 2) {
            npoints = size_t(std::atoll(t[1].c_str()));
            in_cells = false;
        } else if (key == "SCALARS" && t.size() >= 3) {
            const size_t ncomp = t.size() > 3 ? size_t(std::atoi(t[3].c_str())) : 1;
            std::string lt = r.next_nonempty();
            if (upper(lt).rfind("LOOKUP_TABLE", 0) != 0) throw std::runtime_error("SCALARS without LOOKUP_TABLE");
            std::vector<double> v = r.values((in_cells ? ncells : npoints) * ncomp, t[2], binary);
            if (in_cells && t[1] == "wall") {
                g.wall.resize(ncells);
                for (size_t i = 0; i < ncells; ++i) g.wall[i] = int32_t(v[i * ncomp]);
                have_wall = true;
            }
        } else if ((key == "VECTORS" || key == "NORMALS") && t.size() >= 3) {
            r.values((in_cells ? ncells : npoints) * 3, t[2], binary);
        } else if (key == "FIELD" && t.size() >= 3) {
            const int nf = std::atoi(t[2].c_str());
            for (int f = 0; f < nf; ++f) {
                std::vector<std::string> ft = split(r.next_nonempty());
                if (ft.size() < 4) throw std::runtime_error("malformed FIELD array");
                const size_t nc = size_t(std::atoll(ft[1].c_str())), nt = size_t(std::atoll(ft[2].c_str()));
                std::vector<double> v = r.values(nc * nt, ft[3], binary);
                if (in_cells && ft[0] == "wall") {
                    g.wall.resize(nt);
                    for (size_t i = 0; i < nt; ++i) g.wall[i] = int32_t(v[i * nc]);
                    have_wall = true;
                }
            }
        } else if (key == "METADATA") {
            // VTK >= 9 writes a METADATA block (ends at an empty line)
            while (!r.eof() && !r.line().empty()) {}
        }
    }
    for (int i = 0; i < 3; ++i) g.dim[i] = pdim[i] - 1;
    if (xc.size() < 2) throw std::runtime_error("Missing rectilinear x coordinates.");
    for (size_t i = 1; i < xc.size(); ++i) {
        const double sp = std::fabs(xc[i] - xc[i - 1]);
        if (sp > 0.0) { g.dx = sp; break; }
    }
    if (!(g.dx > 0.0)) throw std::runtime_error("Cannot infer positive rectilinear x spacing.");
    if (!have_wall) throw std::runtime_error(file_name + " does not contain a cell-data array named 'wall'.");
    if (int64_t(g.wall.size()) != int64_t(g.dim[0]) * g.dim[1] * g.dim[2])
        throw std::runtime_error(file_name + ": wall array size does not match the grid");
    return g;
}

// ---------------------------------------------------------------------------
// results_<id>/log_rank000 with the reference's time stamp (timeutils.h:6-19)
struct Logger {
    std::string dir;
    void log(const char* fmt, ...) __attribute__((format(printf, 2, 3)))
    {
        char msg[4096];
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(msg, sizeof msg, fmt, ap);
        va_end(ap);
        ::mkdir(dir.c_str(), 0777);
        const std::string fname = dir + "/log_rank000";
        if (FILE* f = std::fopen(fname.c_str(), "at")) {
            const auto now = std::chrono::system_clock::now();
            const std::time_t secs = std::chrono::system_clock::to_time_t(now);
            std::tm date;
            localtime_r(&secs, &date);
            const int ms = int(std::chrono::duration_cast<std::chrono::milliseconds>(
                                   now - std::chrono::system_clock::from_time_t(secs)).count());
            std::fprintf(f, "%04d|%02d|%02d %02d:%02d:%02d.%03d %s\n", date.tm_year + 1900, date.tm_mon + 1,
                         date.tm_mday, date.tm_hour, date.tm_min, date.tm_sec, ms, msg);
            std::fclose(f);
        } else {
            std::printf("unable to create/access file %s", fname.c_str());
        }
        std::printf("%s\n", msg);
        std::fflush(stdout);
    }
};

// output counter: fires when time >= count * period (state.h:66-70)
struct Counter {
    double period = -1.0;
    int count = 0;
    bool action(double t) const { return period > 0 && t >= count * period; }
};

void check(int rc, const char* what)
{
    if (rc != LBM355_OK && rc != LBM355_ETERMINATE)
        throw std::runtime_error(std::string(what) + ": " + lbm355_last_error());
}

int pick(const std::string& v, std::initializer_list<std::pair<const char*, int>> table, const char* what)
{
    for (const auto& e : table)
        if (v == e.first) return e.second;
    throw std::runtime_error(std::string("Unknown ") + what + " '" + v + "'");
}

int run(const Options& opt)
{
    const GeometryData geo = read_geometry(opt.geometry);
    const Material mat = material_by_name(opt.material);
    const double dl = geo.dx;
    if (!(dl > 0.0)) throw std::runtime_error("Geometry spacing must be positive.");
    // physical -> lattice units, sim_tcpc.cu:795-814
    const double dt = opt.lbm_viscosity / mat.nu_inf * dl * dl;
    const double nu0 = mat.carreau_yasuda ? mat.mu0 / mat.density / dl / dl * dt : opt.lbm_viscosity;
    const double lam = mat.carreau_yasuda ? mat.lambda / dt : 1.0;
    const double ivc_flux = opt.ivc_mls * 1e-6 / (dl * dl * dl) * dt;
    const double svc_flux = opt.svc_mls * 1e-6 / (dl * dl * dl) * dt;

    lbm355_case_desc d;
    std::memset(&d, 0, sizeof d);
    d.nx = geo.dim[0]; d.ny = geo.dim[1]; d.nz = geo.dim[2];
    d.geometry = geo.wall.data();
    d.dl = dl; d.dt = dt;
    d.collision = pick(opt.collision, {{"cumulant", LBM355_CUM}, {"cum", LBM355_CUM}, {"bgk", LBM355_BGK},
                                       {"mrt", LBM355_MRT}, {"mrt_les", LBM355_MRT}, {"kbc", LBM355_KBC},
                                       {"kbc_n4", LBM355_KBC}}, "collision");
    if (opt.precision == "bf16")
        throw std::runtime_error("--precision bf16 is not admitted: bfloat16 storage misses FP64 velocity by "
                                 "4e-3 after 40,000 steps of the 512x384x256 junction, above the 1e-3 bar; use fp16");
    if (opt.precision == "fp16" && opt.lbm_viscosity < 1e-2)
        throw std::runtime_error("--precision fp16 is not admitted below --lbm-viscosity 1e-2: half precision does "
                                 "not resolve the non-equilibrium stress there, and the velocity error against FP64 "
                                 "exceeds the 1e-3 bar (TCPC25_03 at 1e-4: 1.6e-3 after 1,000 steps); use fp32");
    d.storage = pick(opt.precision, {{"fp32", LBM355_F32}, {"fp64", LBM355_F64}, {"fp16", LBM355_F16_OFFSET}},
                     "precision");
    d.nu_ref = opt.lbm_viscosity;
    const int cy = pick(opt.cy_model, {{"neq", LBM355_CY_NEQ}, {"off", LBM355_CY_OFF},
                                       {"fd-force", LBM355_CY_FD_FORCE}}, "cy-model");
    d.cy_enabled = mat.carreau_yasuda && cy != LBM355_CY_OFF ? 1 : 0;
    d.cy_model = d.cy_enabled ? cy : LBM355_CY_OFF;
    d.cy_nu0 = nu0; d.cy_lambda = lam; d.cy_a = mat.carreau_yasuda ? mat.a : 2.0;
    d.cy_n = mat.carreau_yasuda ? mat.n : 1.0;
    d.ivc_flux = ivc_flux; d.svc_flux = svc_flux;
    d.profile = pick(opt.profile, {{"plug", LBM355_PROFILE_PLUG}, {"parabolic", LBM355_PROFILE_PARABOLIC}},
                     "inflow-profile");
    d.ivc_center[0] = d.ivc_center[1] = d.svc_center[0] = d.svc_center[1] = -1.0;  // centroid of the plane
    d.pulse_amp = opt.pulse_amp;
    d.pulse_period = opt.pulse_period > 0 ? opt.pulse_period / dt : 0.0;
    d.stats_every = opt.stats_every;
    for (int i = 0; i < 6; ++i) d.face_bc[i] = LBM355_FACE_TCPC;

    Logger L{"results_" + opt.case_id};
    L.log("case=%s material=%s density=%e nu_inf=%e physDl=%e physDt=%e", opt.case_id.c_str(), mat.name.c_str(),
          mat.density, mat.nu_inf, dl, dt);
    L.log("prescribed flows: IVC=%e ml/s SVC=%e ml/s", opt.ivc_mls, opt.svc_mls);

    lbm355_case_info pre;
    int rc = lbm355_case_labels(&d, nullptr, &pre);
    check(rc, "label setup");
    if (d.profile == LBM355_PROFILE_PARABOLIC) {
        // u_max in lattice units; default = twice the plug mean velocity
        auto umax = [&](double phys, double flux, int64_t area) {
            if (phys >= 0) return phys * dt / dl;
            return area > 0 ? 2.0 * flux / double(area) : 0.0;
        };
        d.ivc_umax = umax(opt.ivc_umax, ivc_flux, pre.area_ivc);
        d.svc_umax = umax(opt.svc_umax, svc_flux, pre.area_svc);
    }
    L.log("MPI info: rank=%d, nproc=%d, local_X=%d, offset_X=%d", 0, 1, d.nx, 0);
    // setupBoundaries' log lines, sim_tcpc.cu:632-643 (the label setup runs on the host)
    if (pre.area_ivc <= 0 || pre.area_svc <= 0)
        L.log("error: zero inlet area detected: left=%lld right=%lld", (long long)pre.area_ivc,
              (long long)pre.area_svc);
    if (pre.area_out5 <= 0 || pre.area_out6 <= 0)
        L.log("error: zero outlet area detected: outlet5=%lld outlet6=%lld", (long long)pre.area_out5,
              (long long)pre.area_out6);
    L.log("geometry labels: fluid=%lld wall=%lld IVC(label3)=%lld SVC(label4)=%lld outlet5=%lld outlet6=%lld",
          (long long)pre.fluid, (long long)pre.wall, (long long)pre.area_ivc, (long long)pre.area_svc,
          (long long)pre.area_out5, (long long)pre.area_out6);
    L.log("input VTK cap-label counts before plane imposition: label3=%lld label4=%lld label5=%lld label6=%lld",
          (long long)pre.input_label3, (long long)pre.input_label4, (long long)pre.input_label5,
          (long long)pre.input_label6);

    lbm355_case* c = nullptr;
    rc = lbm355_case_create(&d, &c);
    check(rc, "case setup");
    lbm355_case_info info;
    check(lbm355_case_info_get(c, &info), "case info");
    static const char* ids[4] = {"CUM", "BGK", "MRT_LES", "KBC_N4"};
    L.log("\nSTART: simulation LBM:%s dimensions %d x %d x %d lbmVisc %e physDl %e physDt %e", ids[d.collision],
          d.nx, d.ny, d.nz, opt.lbm_viscosity, dl, dt);
    L.log("lbm355: storage=%s cy_model=%s profile=%s active_cells=%lld bricks=%lld device=%.1f MiB",
          opt.precision.c_str(), d.cy_enabled ? opt.cy_model.c_str() : "off", opt.profile.c_str(),
          (long long)info.active_cells, (long long)info.bricks_stored, double(info.device_bytes) / 1048576.0);
    const bool terminate = info.terminate != 0;
    if (opt.autotune && !terminate) {
        // the launch configuration is timed on a throwaway copy of the case,
        // so the run's own trajectory, counters and outputs are untouched
        lbm355_case* t = nullptr;
        int best = 0;
        if (lbm355_case_create(&d, &t) == LBM355_OK && lbm355_case_autotune(t, &best) == LBM355_OK) {
            const int bpg = best & 0xFF, nt = (best >> 8) & 0xFF, cap = best >> 16;
            check(lbm355_case_set_shape(c, bpg), "launch shape");
            check(lbm355_case_set_nontemporal(c, nt), "cache policy");
            check(lbm355_case_set_occupancy(c, cap), "wavefront cap");
            L.log("lbm355: launch bricks_per_group=%d nontemporal=%d waves_per_cu=%d (timed on a copy)", bpg, nt, cap);
        } else {
            L.log("lbm355: launch autotune skipped (%s)", lbm355_last_error());
        }
        if (t) lbm355_case_destroy(t);
    }

    Counter print_c, vtk_c, reset_c;
    print_c.period = opt.print_period;
    vtk_c.period = opt.vtk_period;
    reset_c.period = opt.stat_reset_period;
    int64_t it = 0, prev_iter = 0, info_iter = 0;
    double start_time = 0.0;
    const std::string state_dir = L.dir + "/current_state";
    auto flag = [&](const char* name) { return L.dir + "/" + name; };
    auto exists = [](const std::string& f) {
        struct stat st;
        return ::stat(f.c_str(), &st) == 0;
    };
    // counter slots of the reference (state.h:72): STAT_RESET 0, PRINT 2, VTK3D 5, SAVESTATE 9
    auto save_state = [&](bool forced) {
        lbm355_state_desc sd;
        std::memset(&sd, 0, sizeof sd);
        sd.phys_final_time = opt.steps > 0 ? double(opt.steps) * dt : opt.final_time;
        for (int k = 0; k < 11; ++k) sd.counter_period[k] = -1.0;
        sd.counter_count[0] = reset_c.count; sd.counter_period[0] = reset_c.period;
        sd.counter_count[2] = print_c.count; sd.counter_period[2] = print_c.period;
        sd.counter_count[5] = vtk_c.count; sd.counter_period[5] = vtk_c.period;
        sd.counter_count[9] = 1;
        sd.stat_prev_iter = prev_iter;
        L.log("[saveState invoked]");
        check(lbm355_case_save_state(c, state_dir.c_str(), &sd), "save state");
        if (forced) {
            if (FILE* f = std::fopen(flag("loadstate").c_str(), "wt")) std::fclose(f);
        }
    };
    if (exists(flag("loadstate"))) {
        // execute(): loadState instead of the initial state (core.h:1511-1517)
        lbm355_state_desc sd;
        std::memset(&sd, 0, sizeof sd);
        L.log("[loadState invoked]");
        check(lbm355_case_load_state(c, state_dir.c_str(), &sd), "load state");
        it = sd.iterations;
        info_iter = it;
        prev_iter = sd.stat_prev_iter;
        reset_c.count = sd.counter_count[0];
        print_c.count = sd.counter_count[2];
        vtk_c.count = sd.counter_count[5];
        start_time = double(it) * dt;
    }
    const auto wall0 = std::chrono::steady_clock::now();
    auto info_clock = wall0;
    const double sites = double(d.nx) * d.ny * d.nz;
    const std::string vtk_dir = L.dir + "/vtk3D";

    // AfterSimUpdate, core.h:1343-1469
    auto after_update = [&]() {
        const double t = double(it) * dt;
        bool write_info = false;
        if (vtk_c.action(t)) {
            ::mkdir(L.dir.c_str(), 0777);
            ::mkdir(vtk_dir.c_str(), 0755);
            char name[512];
            std::snprintf(name, sizeof name, "%s/rank000_data_%d.vtk", vtk_dir.c_str(), vtk_c.count);
            lbm355_output_desc o;
            std::memset(&o, 0, sizeof o);
            o.phys_density = mat.density;
            o.phys_nu_inf = mat.nu_inf;
            o.material_cy = mat.carreau_yasuda ? 1 : 0;
            o.time = t;
            o.cycle = vtk_c.count;
            o.stat_steps = int32_t(it - prev_iter);
            check(lbm355_case_write_vtk(c, name, &o), "vtk output");
            vtk_c.count++;
            write_info = true;
        }
        if (print_c.action(t)) {
            write_info = true;
            print_c.count++;
            if (exists(flag("savestate"))) {
                save_state(false);
                ::remove(flag("savestate").c_str());
                if (FILE* f = std::fopen(flag("savestate_done").c_str(), "wt")) std::fclose(f);
            }
        }
        if (reset_c.action(t)) {
            check(lbm355_case_reset_stats(c), "statistics reset");
            prev_iter = it;
            reset_c.count++;
        }
        if (it > 1 && write_info) {
            check(lbm355_case_sync(c), "sync");
            const auto now = std::chrono::steady_clock::now();
            const double span = std::chrono::duration<double>(now - info_clock).count();
            const double lups = span > 0 ? double(it - info_iter) * sites / span : 0.0;
            const double fluid = span > 0 ? double(it - info_iter) * double(info.active_cells) / span : 0.0;
            info_clock = now;
            info_iter = it;
            const double wt = std::chrono::duration<double>(now - wall0).count();
            const double final_t = opt.steps > 0 ? double(opt.steps) * dt : opt.final_time;
            const double eta = t > start_time ? wt * (final_t - t) / (t - start_time) : 0.0;
            L.log("MLUPS=%.1f iter=%d t=%1.3fs dt=%1.2e lbmVisc=%1.2e WT=%.0fs ETA=%.0fs", lups * 1e-6, int(it), t,
                  dt, opt.lbm_viscosity, wt, eta);
            double flux[4];
            check(lbm355_case_flux(c, 0, flux), "flux");
            const double to_lmin = dl * dl * dl / dt * 6.0e4;
            L.log("lbm355: fluid MLUPS=%.1f flux l/min IVC=%.6f SVC=%.6f RPA(5)=%.6f LPA(6)=%.6f", fluid * 1e-6,
                  flux[0] * to_lmin, flux[1] * to_lmin, flux[2] * to_lmin, flux[3] * to_lmin);
        }
    };
    auto any_action = [&](int64_t k) {
        const double t = double(k) * dt;
        return vtk_c.action(t) || print_c.action(t) || reset_c.action(t);
    };
    auto finished = [&](int64_t k) {
        return opt.steps > 0 ? k >= opt.steps : double(k) * dt > opt.final_time;
    };

    check(lbm355_case_set_inflow(c, ivc_flux, svc_flux), "inflow");  // updateKernelVelocities
    after_update();
    int code = 0;
    for (;;) {
        // step up to the next iteration that fires a counter or ends the run;
        // a terminated case stops after its first step like the reference
        int64_t k = it + 1;
        while (!terminate && !any_action(k) && !finished(k) && k - it < 4096) ++k;
        check(lbm355_case_step(c, k - it), "step");
        it = k;
        after_update();
        if (opt.max_wall_time > 0 &&
            std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count() >= opt.max_wall_time) {
            // wallTimeReached: save the state, flag it for loading, stop (core.h:1580-1591)
            check(lbm355_case_sync(c), "sync");
            L.log("maximum wall time reached");
            save_state(true);
            break;
        }
        if (finished(it)) {
            L.log(opt.steps > 0 ? "requested steps reached" : "physFinalTime reached");
            if (opt.save_at_end) save_state(true);
            break;
        }
        if (terminate) {
            L.log("terminate flag triggered");
            code = 2;
            break;
        }
    }
    check(lbm355_case_sync(c), "sync");
    lbm355_case_destroy(c);
    return code;
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        const Options options = parse_args(argc, argv);
        return run(options);
    } catch (const std::exception& e) {
        std::cerr << "sim_tcpc: " << e.what() << "\n";
        usage();
        return 1;
    }
}
