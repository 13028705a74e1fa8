#include "Case.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

#include "ConfigReader.h"

static int axis_of(const std::string& s) {
    if (s == "x" || s == "0") return 0;
    if (s == "y" || s == "1") return 1;
    if (s == "z" || s == "2") return 2;
    return -1;
}

// The plane value is parsed as long double (the reference compares against
// long double literals, ComplexGeometry.cpp:586-632) and split into hi + lo.
static void split_ld(const std::string& text, double& hi, double& lo) {
    const long double v = std::strtold(text.c_str(), nullptr);
    hi = (double)v;
    lo = (double)(v - (long double)hi);
}

bool load_case(const std::string& path, CaseConfig& c, std::string& err) {
    ConfigReader reader(path);
    std::map<std::string, double> num;
    std::multimap<std::string, std::string> str;
    if (!reader.readConfigData(num) || !reader.readConfigStrings(str)) {
        err = reader.error();
        return false;
    }
    auto get = [&](const char* k, double def) { auto it = num.find(k); return it == num.end() ? def : it->second; };
    auto getstr = [&](const char* k) -> std::string { auto it = str.find(k); return it == str.end() ? "" : it->second; };
    c.stl_file = getstr("stl_file");
    if (!c.stl_file.empty() && c.stl_file[0] != '/') {
        const size_t slash = path.find_last_of('/');
        if (slash != std::string::npos) c.stl_file = path.substr(0, slash + 1) + c.stl_file;
    }
    c.spacing = get("spacing", c.spacing);
    c.q = (int)get("lattice", 19);
    c.quirks = (int)get("quirks", 1);
    c.tau = get("tau", c.tau);
    c.steps = (long long)get("steps", (double)c.steps);
    c.output_every = (int)get("output_every", c.output_every);
    c.device = (int)get("device", 0);
    if (!getstr("output_dir").empty()) c.output_dir = getstr("output_dir");
    c.autotune = get("autotune", 0) != 0;
    c.boundary_points = get("boundary_points", 0) != 0;
    const std::string of = getstr("output_format");
    if (of == "raw") c.output_raw = true;
    else if (!of.empty() && of != "ascii") { err = "output_format must be ascii or raw"; return false; }
    c.autotune_cache = getstr("autotune_cache");
    const std::string st = getstr("storage");
    if (st == "f32") c.storage = LBM_STORE_F32;
    else if (st == "f16s" || st == "f16") c.storage = LBM_STORE_F16S;
    else if (!st.empty() && st != "f64") { err = "unknown storage " + st; return false; }
    const std::string sc = getstr("scheme");
    if (sc == "aa") c.scheme = LBM_SCHEME_AA;
    else if (!sc.empty() && sc != "two_lattice") { err = "unknown scheme " + sc; return false; }
    if (!getstr("box_dims").empty()) {
        std::istringstream a(getstr("box_origin")), b(getstr("box_dims"));
        a >> c.box_origin[0] >> c.box_origin[1] >> c.box_origin[2];
        b >> c.box_dims[0] >> c.box_dims[1] >> c.box_dims[2];
        c.explicit_box = true;
    }
    // planes keep the order in which they appear in the file: read the file again in order
    std::map<std::string, int> index;
    {
        std::ifstream in(path);
        std::string line;
        while (std::getline(in, line)) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream ls(line);
            std::string key;
            ls >> key;
            if (key == "plane") {
                std::string name, ax, val;
                ls >> name >> ax >> val;
                lbm_plane_desc p;
                std::memset(&p, 0, sizeof p);
                p.axis = axis_of(ax);
                if (name.empty() || p.axis < 0 || val.empty()) { err = "bad plane line: " + line; return false; }
                split_ld(val, p.value_hi, p.value_lo);
                p.tol = 1e-2;  // ComplexGeometry.cpp:580
                std::string w;
                while (ls >> w) {
                    if (w == "tol") ls >> p.tol;
                    else if (w == "range" && p.nranges < 3) {
                        std::string rax;
                        ls >> rax >> p.range_lo[p.nranges] >> p.range_hi[p.nranges];
                        p.range_axis[p.nranges] = axis_of(rax);
                        if (p.range_axis[p.nranges] < 0) { err = "bad range axis: " + line; return false; }
                        p.nranges++;
                    } else { err = "bad plane option '" + w + "': " + line; return false; }
                }
                index[name] = (int)c.planes.size();
                c.plane_names.push_back(name);
                c.planes.push_back(p);
            }
        }
    }
    c.bcs.assign(c.planes.size(), lbm_bc_desc());
    std::vector<bool> seen(c.planes.size(), false);
    for (auto it = str.lower_bound("bc"); it != str.upper_bound("bc"); ++it) {
        std::istringstream ls(it->second);
        std::string name, kind;
        int sigma = 0;
        ls >> name >> kind >> sigma;
        auto pi = index.find(name);
        if (pi == index.end()) { err = "bc for unknown plane " + name; return false; }
        lbm_bc_desc b;
        std::memset(&b, 0, sizeof b);
        b.axis = c.planes[pi->second].axis;
        b.sigma = sigma;
        b.waveform = -1;
        b.radius = 1.0;
        b.rho = 1.0;
        if (kind == "velocity") {
            b.kind = LBM_BC_VELOCITY;
            ls >> b.u[0] >> b.u[1] >> b.u[2];
        } else if (kind == "pressure") {
            b.kind = LBM_BC_PRESSURE;
            ls >> b.rho;
        } else { err = "bc kind must be velocity or pressure: " + it->second; return false; }
        std::string w;
        while (ls >> w) {
            if (w == "parabolic") { b.profile = LBM_PROFILE_PARABOLIC; ls >> b.center[0] >> b.center[1] >> b.radius; }
            else if (w == "waveform") ls >> b.waveform;
            else if (w == "typo") b.typo = 1;
            else { err = "bad bc option '" + w + "'"; return false; }
        }
        c.bcs[pi->second] = b;
        seen[pi->second] = true;
    }
    for (size_t p = 0; p < seen.size(); ++p)
        if (!seen[p]) { err = "plane " + c.plane_names[p] + " has no bc line"; return false; }
    for (auto it = str.lower_bound("waveform"); it != str.upper_bound("waveform"); ++it) {
        std::istringstream ls(it->second);
        WaveSpec w;
        std::string kind;
        ls >> w.slot >> kind;
        if (kind == "fourier") {
            w.kind = LBM_WAVE_FOURIER;
            ls >> w.period;
            double a, ph;
            while (ls >> a >> ph) { w.a.push_back(a); w.phi.push_back(ph); }
        } else if (kind == "pv2013") {
            w.kind = LBM_WAVE_PV2013;
            w.period = 0.9;  // pulsatile.h:10
            ls >> w.dt;
        } else { err = "waveform kind must be fourier or pv2013"; return false; }
        c.waves.push_back(w);
    }
    if (c.stl_file.empty() && !c.explicit_box) { err = "stl_file missing"; return false; }
    return true;
}
