// red_driver.cpp -- drop-in replacement for the reference's `red` executable main loop.
//
// Same command line (red.cuda/src/red.cuda/options.cpp:59-195), same parameter file
// (parameter.cpp:13-292), body file (pp_disk.cu:1307-1322, 1576-1612), output files and
// names (main.cu:28-71), result/event/log/info layouts (pp_disk.cu:1685-1848,
// main.cu:73-112, file_util.cpp:213-267) and the driver loop of main.cu:129-311.  The
// integration itself runs on the GPU through the C ABI in include/rednative.h.
//
// Additions: two optional parameter keys, `dt` (initial step in days; the reference
// hard-codes 0.001 in main.cu:141) and `encounter_hill_factor` (close encounters at that
// many mutual Hill radii), the `-mixed` switch (mixed-precision force sweeps) and
// `-restart <result file>`: the run resumes from the last snapshot of a text result file.
// The reference has no explicit resume; its text snapshot has the body file's 16 columns
// with t in the epoch column (pp_disk.cu:1707-1722 against 1576-1612), so "last snapshot
// + counts line + start_time" is its implicit restart, which -restart performs: bodies and
// t from the snapshot, the stop time of the parameter file's schedule kept.  The step
// controller's state is not in the snapshot (nor in the reference's): the first step is
// the `dt` parameter.
// `-cpu` is refused: this build has one execution path, the GPU one.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/rednative.h"

namespace {

const double GAUSS = 1.720209895e-2;     // red_constants.h:21
const double YEAR_TO_DAY = 365.25;       // red_constants.h:75
const char SEP = ' ';

struct Error {
    std::string msg;
};

std::string time_stamp()  // tools.cpp:76-84
{
    char buf[20];
    time_t now = time(nullptr);
    strftime(buf, 20, "%Y-%m-%d %H:%M:%S", localtime(&now));
    return buf;
}

bool is_number(const std::string& s)  // tools.cpp:20-34
{
    for (char c : s)
        if (!(std::isdigit((unsigned char)c) || c == 'e' || c == 'E' || c == '.' || c == '-' || c == '+')) return false;
    return true;
}

void trim(std::string& s)  // tools.cpp:37-74
{
    size_t e = s.find_last_not_of(" \t");
    if (e != std::string::npos) s = s.substr(0, e + 1);
    size_t b = s.find_first_not_of(" \t");
    if (b != std::string::npos) s = s.substr(b);
}

std::string combine_path(const std::string& dir, const std::string& file)  // file_util.cpp
{
    if (dir.empty()) return file;
    if (dir.back() == '/' || dir.back() == '\\') return dir + file;
    return dir + "/" + file;
}

// ---- options (options.cpp:59-195) -------------------------------------------------------
struct Options {
    int id_dev = 0;
    bool gpu = true, verbose = false, ups = false, ef = false, mixed = false, stepwise = false;
    int n_tpb = 64;
    std::string input_dir, printout_dir, param_file, gasdisk_file, body_file, restart;
    std::string info = "info", event = "event", log = "log", result = "result";
};

void usage()
{
    std::cout << "Usage: red <parameterlist>\n"
                 "     -id_dev | --id_active_device             : the id of the device which will execute the code\n"
                 "     -ups    | --use-padded-storage           : the padded kernel's rules: inactive bodies keep\n"
                 "                                                pulling and being pulled until the next rebuild\n"
                 "     -n_tpb  | --n_thread-per-block           : threads per block of the force sweeps\n"
                 "     -iDir   | --inputDir <directory>         : the directory containing the input files\n"
                 "     -p      | --parameter <filename>         : the file containing the parameters of the simulation\n"
                 "     -gd     | --gas_disk <filename>          : accepted; the gas disk is not applied (as in the reference)\n"
                 "     -ic     | --initial_condition <filename> : the file containing the initial conditions\n"
                 "     -info   | --info-filename                : info file name (default info)\n"
                 "     -event  | --event-filename               : event file name (default event)\n"
                 "     -log    | --log-filename                 : log file name (default log)\n"
                 "     -result | --result-filename              : result file name (default result)\n"
                 "     -v      | --verbose                      : verbose mode\n"
                 "     -ef                                      : prefix output names with the run configuration\n"
                 "     -mixed                                   : mixed-precision force sweeps\n"
                 "     -stepwise                                : fixed-step runs step from the host one at a time\n"
                 "     -restart <filename>                      : resume from the last snapshot of a text result file\n"
                 "     -gpu    | --gpu                          : execute on the GPU (the only path of this build)\n"
                 "     -h      | --help                         : print this help\n";
}

Options parse_options(int argc, const char** argv)
{
    Options o;
    auto next = [&](int& i, const std::string& p) -> std::string {
        if (i + 1 >= argc) throw Error{"Missing value at: " + p};
        return argv[++i];
    };
    for (int i = 1; i < argc; i++) {
        std::string p = argv[i];
        if (p == "--id_active_device" || p == "-id_dev") {
            std::string v = next(i, p);
            if (!is_number(v)) throw Error{"Invalid number at: " + p};
            o.id_dev = atoi(v.c_str());
        } else if (p == "--use-padded-storage" || p == "-ups") {
            o.ups = true;
        } else if (p == "--n_thread-per-block" || p == "-n_tpb") {
            std::string v = next(i, p);
            if (!is_number(v)) throw Error{"Invalid number at: " + p};
            o.n_tpb = atoi(v.c_str());
        } else if (p == "--inputDir" || p == "-iDir") {
            o.input_dir = next(i, p);
            o.printout_dir = o.input_dir;
        } else if (p == "--parameters" || p == "-p") {
            o.param_file = next(i, p);
        } else if (p == "--gas_disk" || p == "-gd") {
            o.gasdisk_file = next(i, p);
        } else if (p == "--initial_conditions" || p == "-ic") {
            o.body_file = next(i, p);
        } else if (p == "--info-filename" || p == "-info") {
            o.info = next(i, p);
        } else if (p == "--event-filename" || p == "-event") {
            o.event = next(i, p);
        } else if (p == "--log-filename" || p == "-log") {
            o.log = next(i, p);
        } else if (p == "--result-filename" || p == "-result") {
            o.result = next(i, p);
        } else if (p == "--verbose" || p == "-v") {
            o.verbose = true;
        } else if (p == "-ef") {
            o.ef = true;
        } else if (p == "-mixed") {
            o.mixed = true;
        } else if (p == "-stepwise") {
            o.stepwise = true;
        } else if (p == "-restart") {
            o.restart = next(i, p);
        } else if (p == "--cpu" || p == "-cpu") {
            throw Error{"-cpu: this build integrates on the GPU only"};
        } else if (p == "--gpu" || p == "-gpu") {
            o.gpu = true;
        } else if (p == "--help" || p == "-h") {
            usage();
            exit(EXIT_SUCCESS);
        } else {
            throw Error{"Invalid switch on command-line: " + p + "."};
        }
    }
    if (o.param_file.empty()) throw Error{"Missing filename for -p | --parameter!"};
    if (o.body_file.empty() && o.restart.empty()) throw Error{"Missing filename for -ic | --initial_conditions!"};
    return o;
}

// ---- parameter file (parameter.cpp:13-292, file_util.cpp:85-108) --------------------------
struct Params {
    std::string name, description;
    int integrator = RED_EULER;
    bool adaptive = false, error_check_for_tp = false, close_encounter = false, inner_steps = true;
    double tolerance = 1.0e-10, start_time = 0.0, length = 0.0, output_interval = 0.0, stop_time = 0.0;
    bool binary = false;
    int frame_center = RED_FRAME_BARY;
    double hit_centrum = 0.0, ejection = 0.0, radii_enhance_factor = 0.0, hill = 0.0, dt = 0.001;
};

Params load_parameters(const std::string& path, bool verbose)
{
    std::ifstream f(path.c_str());
    if (!f) throw Error{"The file '" + path + "' could not opened!"};
    Params P;
    std::string line;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty() || line[0] == '#') continue;
        size_t h = line.find('#');
        if (h != std::string::npos) line = line.substr(0, h);
        // tokenizer on '=' skips empty tokens (tokenizer.cpp)
        std::vector<std::string> tok;
        std::string cur;
        for (char c : line) {
            if (c == '=') { if (!cur.empty()) tok.push_back(cur); cur.clear(); }
            else cur.push_back(c);
        }
        if (!cur.empty()) tok.push_back(cur);
        if (tok.empty()) continue;
        if (tok.size() < 2) throw Error{"Invalid key/value pair: " + line + "."};
        std::string key = tok[0], value = tok[1];
        trim(key);
        trim(value);
        std::transform(key.begin(), key.end(), key.begin(), ::tolower);
        auto num = [&]() {
            if (!is_number(value)) throw Error{"Invalid number at: " + key};
            return atof(value.c_str());
        };
        auto boolean = [&]() {
            if (value == "true") return true;
            if (value == "false") return false;
            throw Error{"Invalid value at: " + key};
        };
        if (key == "name") P.name = value;
        else if (key == "description") P.description = value;
        else if (key == "integrator") {
            std::string v = value;
            std::transform(v.begin(), v.end(), v.begin(), ::tolower);
            if (v == "e" || v == "euler") P.integrator = RED_EULER;
            else if (v == "rk2" || v == "rungekutta2") P.integrator = RED_RK2;
            else if (v == "rk4" || v == "rungekutta4") P.integrator = RED_RK4;
            else if (v == "rk5" || v == "rungekutta5") P.integrator = RED_RK5;
            else if (v == "rkf8" || v == "rungekuttafehlberg8") P.integrator = RED_RKF78;
            else if (v == "rkn" || v == "rungekuttanystrom") throw Error{"Requested integrator is not implemented."};
            else throw Error{"Invalid integrator type: " + v};
        } else if (key == "tolerance") { P.tolerance = num(); P.adaptive = true; }
        else if (key == "error_check_for_tp") P.error_check_for_tp = boolean();
        else if (key == "close_encounter") P.close_encounter = boolean();
        else if (key == "inner_steps") P.inner_steps = boolean();
        else if (key == "start_time") P.start_time = num() * YEAR_TO_DAY;
        else if (key == "length") P.length = num() * YEAR_TO_DAY;
        else if (key == "output_interval") P.output_interval = num() * YEAR_TO_DAY;
        else if (key == "ejection") P.ejection = num();
        else if (key == "hit_centrum") P.hit_centrum = num();
        else if (key == "radii_enhance_factor" || key == "collision_factor") P.radii_enhance_factor = num();
        else if (key == "output_type") {
            if (value == "text" || value == "ascii") P.binary = false;
            else if (value == "binary") P.binary = true;
            else throw Error{"Invalid value at: " + key};
        } else if (key == "frame_center") {
            if (value == "astro") P.frame_center = RED_FRAME_ASTRO;
            else if (value == "bary") P.frame_center = RED_FRAME_BARY;
            else throw Error{"Invalid value at: " + key};
        } else if (key == "dt") P.dt = num();
        else if (key == "encounter_hill_factor") P.hill = num();
        else throw Error{"Invalid parameter: " + key + "."};
        if (verbose) std::cout << "\t'" << key << "' was assigned to '" << value << "'" << std::endl;
    }
    P.start_time *= GAUSS;   // parameter.cpp:287-292
    P.length *= GAUSS;
    P.output_interval *= GAUSS;
    P.stop_time = P.start_time + P.length;
    return P;
}

// ---- body file (pp_disk.cu:1307-1322, 1576-1612) -----------------------------------------
struct Bodies {
    int counts[7] = {0};
    std::vector<int> id, type, mig_type;
    std::vector<std::string> names;
    std::vector<double> epoch, mass, radius, density, cd, mig_stop, r, v;
    int n() const { int s = 0; for (int k = 0; k < 7; k++) s += counts[k]; return s; }
};

Bodies load_bodies(const std::string& path)
{
    std::ifstream in(path.c_str());
    if (!in) throw Error{"Cannot open " + path + "."};
    Bodies B;
    for (int k = 0; k < 7; k++) in >> B.counts[k];
    const int n = B.n();
    B.id.resize(n); B.type.resize(n); B.mig_type.resize(n); B.names.resize(n);
    B.epoch.resize(n); B.mass.resize(n); B.radius.resize(n); B.density.resize(n); B.cd.resize(n);
    B.mig_stop.resize(n); B.r.resize(3 * n); B.v.resize(3 * n);
    for (int k = 0; k < n; k++) {
        std::string name;
        in >> B.id[k] >> name >> B.type[k] >> B.epoch[k] >> B.mass[k] >> B.radius[k] >> B.density[k] >> B.cd[k] >>
            B.mig_type[k] >> B.mig_stop[k] >> B.r[3 * k] >> B.r[3 * k + 1] >> B.r[3 * k + 2] >> B.v[3 * k] >>
            B.v[3 * k + 1] >> B.v[3 * k + 2];
        if (!in) throw Error{"Cannot read body record " + std::to_string(k) + " of " + path};
        if (name.length() > 30) name = name.substr(0, 30);
        B.names[k] = name;
    }
    return B;
}

// Last snapshot of a text result file (pp_disk.cu:1707-1722): the trailing lines that share
// the time column.  Returns the bodies (counts from the body types, which must be in the
// body file's type order) and the snapshot time in days.
Bodies load_restart(const std::string& path, double* t_day)
{
    std::ifstream in(path.c_str());
    if (!in) throw Error{"Cannot open " + path + "."};
    std::vector<std::vector<std::string>> rows;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::vector<std::string> tok;
        std::string w;
        while (ls >> w) tok.push_back(w);
        if (tok.empty()) continue;
        if (tok.size() != 16) throw Error{"Invalid result line in " + path + ": " + line};
        rows.push_back(tok);
    }
    if (rows.empty()) throw Error{"No snapshot in " + path + "."};
    const double t = atof(rows.back()[3].c_str());
    size_t first = rows.size();
    while (first > 0 && atof(rows[first - 1][3].c_str()) == t) first--;
    Bodies B;
    int last_type = 0;
    for (size_t k = first; k < rows.size(); k++) {
        const auto& x = rows[k];
        auto f = [&](int c) { return atof(x[c].c_str()); };
        const int type = atoi(x[2].c_str());
        if (type < 0 || type > 6 || type < last_type) throw Error{"Snapshot bodies out of type order in " + path + "."};
        last_type = type;
        B.counts[type]++;
        B.id.push_back(atoi(x[0].c_str()));
        B.names.push_back(x[1].substr(0, 30));
        B.type.push_back(type);
        B.epoch.push_back(f(3));
        B.mass.push_back(f(4)); B.radius.push_back(f(5)); B.density.push_back(f(6)); B.cd.push_back(f(7));
        B.mig_type.push_back(atoi(x[8].c_str()));
        B.mig_stop.push_back(f(9));
        for (int c = 0; c < 3; c++) B.r.push_back(f(10 + c));
        for (int c = 0; c < 3; c++) B.v.push_back(f(13 + c));
    }
    *t_day = t;
    return B;
}

// ---- output -------------------------------------------------------------------------------
struct Sim {
    red_ctx* ctx = nullptr;
    std::vector<std::string> names;   // kept aligned with the context's body order
    std::vector<int> mig_type;
    std::vector<double> mig_stop;
};

void check(red_ctx* ctx, int rc)
{
    if (rc != RED_OK) throw Error{std::string("rednative error ") + std::to_string(rc) + ": " + red_last_error(ctx)};
}

// pp_disk.cu:1685-1725
void print_result_ascii(std::ostream& sout, Sim& S)
{
    int cnt[7];
    check(S.ctx, red_get_counts(S.ctx, cnt));
    int n = 0;
    for (int k = 0; k < 7; k++) n += cnt[k];
    std::vector<double> r(3 * n), v(3 * n), m(n), R(n), rho(n), cd(n);
    std::vector<int> id(n), type(n);
    check(S.ctx, red_get_state(S.ctx, r.data(), v.data(), m.data(), R.data(), rho.data(), cd.data(), id.data(), type.data()));
    double t;
    check(S.ctx, red_get_time(S.ctx, &t));
    sout.precision(16);
    sout.setf(std::ios::right);
    sout.setf(std::ios::scientific);
    const int iw = 8, vw = 25;
    for (int i = 0; i < n; i++) {
        if (id[i] <= 0 || type[i] >= 7) continue;
        sout << std::setw(iw) << id[i] << SEP << std::setw(30) << S.names[i] << SEP << std::setw(2) << type[i] << SEP
             << std::setw(vw) << t / GAUSS << SEP << std::setw(vw) << m[i] << SEP << std::setw(vw) << R[i] << SEP
             << std::setw(vw) << rho[i] << SEP << std::setw(vw) << cd[i] << SEP << std::setw(2) << S.mig_type[i] << SEP
             << std::setw(vw) << S.mig_stop[i] << SEP << std::setw(vw) << r[3 * i] << SEP << std::setw(vw) << r[3 * i + 1]
             << SEP << std::setw(vw) << r[3 * i + 2] << SEP << std::setw(vw) << v[3 * i] * GAUSS << SEP << std::setw(vw)
             << v[3 * i + 1] * GAUSS << SEP << std::setw(vw) << v[3 * i + 2] * GAUSS << std::endl;
    }
    sout.flush();
}

// pp_disk.cu:1727-1770
void print_result_binary(std::ostream& sout, Sim& S)
{
    int cnt[7];
    check(S.ctx, red_get_counts(S.ctx, cnt));
    int n = 0;
    for (int k = 0; k < 7; k++) n += cnt[k];
    std::vector<double> r(3 * n), v(3 * n), m(n), R(n), rho(n), cd(n);
    std::vector<int> id(n), type(n);
    check(S.ctx, red_get_state(S.ctx, r.data(), v.data(), m.data(), R.data(), rho.data(), cd.data(), id.data(), type.data()));
    double t;
    check(S.ctx, red_get_time(S.ctx, &t));
    double tout = t / GAUSS;
    sout.write(reinterpret_cast<const char*>(&tout), sizeof(tout));
    for (int i = 0; i < n; i++) {
        if (id[i] <= 0 || type[i] >= 7) continue;
        double vo[3] = {v[3 * i] * GAUSS, v[3 * i + 1] * GAUSS, v[3 * i + 2] * GAUSS};
        sout.write(reinterpret_cast<const char*>(&id[i]), 4);
        sout.write(S.names[i].c_str(), S.names[i].size() + 1);
        sout.write(reinterpret_cast<const char*>(&type[i]), 4);
        sout.write(reinterpret_cast<const char*>(&m[i]), 8);
        sout.write(reinterpret_cast<const char*>(&R[i]), 8);
        sout.write(reinterpret_cast<const char*>(&rho[i]), 8);
        sout.write(reinterpret_cast<const char*>(&cd[i]), 8);
        sout.write(reinterpret_cast<const char*>(&S.mig_type[i]), 4);
        sout.write(reinterpret_cast<const char*>(&S.mig_stop[i]), 8);
        sout.write(reinterpret_cast<const char*>(&r[3 * i]), 24);
        sout.write(reinterpret_cast<const char*>(vo), 24);
    }
    sout.flush();
}

void print_result(std::ostream& sout, Sim& S, bool binary)
{
    if (binary) print_result_binary(sout, S);
    else print_result_ascii(sout, S);
}

// pp_disk.cu:1772-1848
void print_event_data(std::ostream& sout, std::ostream& log_f, Sim& S)
{
    static const char* e_names[] = {"NONE", "HIT_CENTRUM", "EJECTION", "CLOSE_ENCOUNTER", "COLLISION"};
    int n = 0;
    check(S.ctx, red_sp_events(S.ctx, nullptr, 0, &n));
    std::vector<red_event> ev(std::max(n, 1));
    check(S.ctx, red_sp_events(S.ctx, ev.data(), n, &n));
    sout.precision(16);
    sout.setf(std::ios::right);
    sout.setf(std::ios::scientific);
    log_f.precision(16);
    log_f.setf(std::ios::right);
    log_f.setf(std::ios::scientific);
    const int iw = 8, vw = 25;
    for (int k = 0; k < n; k++) {
        const red_event& e = ev[k];
        sout << std::setw(16) << e_names[e.name] << SEP << std::setw(vw) << e.t / GAUSS << SEP << std::setw(vw) << e.d
             << SEP << std::setw(iw) << e.id1 << SEP << std::setw(iw) << e.id2 << SEP;
        auto body = [&](const double* p, const double* r, const double* v, const double* oe) {
            sout << std::setw(vw) << p[0] << SEP << std::setw(vw) << p[2] << SEP << std::setw(vw) << p[1] << SEP;
            for (int c = 0; c < 3; c++) sout << std::setw(vw) << r[c] << SEP;
            for (int c = 0; c < 3; c++) sout << std::setw(vw) << v[c] * GAUSS << SEP;
            for (int c = 0; c < 6; c++) sout << std::setw(vw) << oe[c] << SEP;
        };
        body(e.p1, e.r1, e.v1, e.oe1);
        body(e.p2, e.r2, e.v2, e.oe2);
        body(e.ps, e.rs, e.vs, e.oes);
        sout << std::endl;
        if (log_f)
            log_f << time_stamp() << SEP << e_names[e.name] << SEP << std::setw(iw) << e.id1 << SEP << std::setw(iw)
                  << e.id2 << SEP << std::endl;
    }
    sout.flush();
    log_f.flush();
}

// main.cu:73-112
void print_info(std::ostream& sout, Sim& S, double dt, clock_t sum_steps, clock_t one_step, clock_t* info_start)
{
    int c[8], cnt[7];
    check(S.ctx, red_counters(S.ctx, c));
    check(S.ctx, red_get_counts(S.ctx, cnt));
    double t;
    check(S.ctx, red_get_time(S.ctx, &t));
    int n = 0;
    for (int k = 0; k < 7; k++) n += cnt[k];
    *info_start = clock();
    for (std::ostream* o : {&std::cout, &sout}) {
        o->setf(std::ios::right);
        o->setf(std::ios::scientific);
        *o << time_stamp() << " t: " << std::setprecision(6) << std::setw(12) << t / GAUSS << ", dt: "
           << std::setprecision(6) << std::setw(12) << dt / GAUSS;
        *o << ", dT_cpu: " << std::setprecision(3) << std::setw(10) << one_step / (double)CLOCKS_PER_SEC << " s";
        *o << ", dT_avg: " << std::setprecision(3) << std::setw(10)
           << (sum_steps / (double)CLOCKS_PER_SEC) / std::max(1, c[2]) << " s";
        *o << ", Nc: " << std::setw(5) << c[3] << ", Nce: " << std::setw(5) << c[6] << ", Ne: " << std::setw(5) << c[4]
           << ", Nh: " << std::setw(5) << c[5] << ", N : " << std::setw(6) << n << std::endl;
    }
}

}  // namespace

int main(int argc, const char** argv, const char** env)
{
    time_t start = time(nullptr);
    std::ofstream result_f, info_f, event_f, log_f;
    Sim S;
    int exit_code = EXIT_SUCCESS;
    try {
        Options opt = parse_options(argc, argv);
        Params P = load_parameters(combine_path(opt.input_dir, opt.param_file), opt.verbose);
        double restart_day = 0.0;
        Bodies B = opt.restart.empty() ? load_bodies(combine_path(opt.input_dir, opt.body_file))
                                       : load_restart(combine_path(opt.input_dir, opt.restart), &restart_day);
        if (!opt.restart.empty()) P.start_time = restart_day * GAUSS;   // stop_time keeps the schedule
        const int n = B.n();
        // options::create_pp_disk (options.cpp:197-235): barycentric frame when the star rests
        // at the origin, then epochs * k and v / k
        if (n > 0 && B.r[0] == 0.0 && B.r[1] == 0.0 && B.r[2] == 0.0 && B.v[0] == 0.0 && B.v[1] == 0.0 && B.v[2] == 0.0) {
            if (P.frame_center == RED_FRAME_BARY) red_host_transform_to_bc(n, B.mass.data(), B.r.data(), B.v.data());
        } else if (P.frame_center == RED_FRAME_ASTRO) {
            if (red_host_transform_to_ac(n, B.type.data(), B.r.data(), B.v.data()) != RED_OK)
                throw Error{"No star is included!"};
        }
        for (auto& e : B.epoch) e *= GAUSS;
        red_host_transform_velocity(n, B.v.data());

        red_bodies rb;
        memcpy(rb.counts, B.counts, sizeof(rb.counts));
        rb.id = B.id.data(); rb.type = B.type.data(); rb.mig_type = B.mig_type.data();
        rb.mig_stop_at = B.mig_stop.data(); rb.epoch = B.epoch.data(); rb.mass = B.mass.data();
        rb.radius = B.radius.data(); rb.density = B.density.data(); rb.cd = B.cd.data();
        rb.r = B.r.data(); rb.v = B.v.data();
        red_params rp;
        memset(&rp, 0, sizeof(rp));
        rp.integrator = P.integrator;
        rp.adaptive = P.adaptive;
        rp.tolerance = P.tolerance;
        rp.error_check_for_tp = P.error_check_for_tp;
        rp.close_encounter = P.close_encounter;
        rp.inner_steps = P.inner_steps;
        rp.frame_center = P.frame_center;
        rp.hit_centrum = P.hit_centrum;
        rp.ejection = P.ejection;
        rp.radii_enhance_factor = P.radii_enhance_factor;
        rp.encounter_hill_factor = P.hill;
        rp.dt_day = P.dt;
        rp.precision = opt.mixed ? RED_PREC_MIXED : RED_PREC_FP64;
        rp.padded_storage = opt.ups ? 1 : 0;
        rp.device = opt.id_dev;
        rp.block = opt.n_tpb >= 64 ? opt.n_tpb : 0;
        int rc = red_ctx_create(&rb, &rp, &S.ctx);
        if (rc != RED_OK) throw Error{std::string("red_ctx_create: ") + red_last_error(S.ctx)};
        S.names = B.names;
        S.mig_type = B.mig_type;
        S.mig_stop = B.mig_stop;
        check(S.ctx, red_set_time(S.ctx, P.start_time));

        // open_streams (main.cu:28-71)
        std::string prefix;
        if (opt.ef) {
            const char* short_names[] = {"E", "RK2", "RK4", "RK5", "RKF8"};
            prefix = std::string("R_gpu_") + (opt.ups ? "ps" : "ns") + "_" + (P.adaptive ? "as" : "fs") + "_" +
                     short_names[P.integrator] + "_";
        }
        const std::string dir = opt.printout_dir;
        result_f.open(combine_path(dir, prefix + opt.result) + (P.binary ? ".bin" : ".txt"),
                      P.binary ? (std::ios::out | std::ios::binary) : std::ios::out);
        info_f.open(combine_path(dir, prefix + opt.info) + ".txt");
        event_f.open(combine_path(dir, prefix + opt.event) + ".txt");
        log_f.open(combine_path(dir, prefix + opt.log) + ".txt");

        // file_util.cpp:213-257
        log_f << time_stamp() << " starting " << argv[0] << std::endl << "Command line arguments: " << std::endl;
        for (int i = 1; i < argc; i++) log_f << argv[i] << SEP;
        log_f << std::endl << std::endl;
        for (const char** e = env; e && *e; e++) {
            std::string s = *e;
            if (s.find("HOSTNAME=") < s.length() || s.find("USER=") < s.length() || s.find("OSTYPE=") < s.length())
                log_f << s << std::endl;
        }
        log_f << std::endl;

        // main loop (main.cu:159-263)
        double ps = 0.0, dt = 0.0;
        clock_t sum_steps = 0, one_step = 0, info_start = clock();
        print_result(result_f, S, P.binary);
        double t = P.start_time;
        const bool tests = 0.0 < P.radii_enhance_factor || 0.0 < P.hill;
        // the checks after a step (main.cu:225-255)
        auto after_step = [&]() {
            int handled = 0;
            if (P.close_encounter && P.inner_steps) {
                if (tests) {
                    check(S.ctx, red_check_close_encounter(S.ctx, &handled));
                    if (handled) {
                        print_event_data(event_f, log_f, S);
                        check(S.ctx, red_clear_events(S.ctx));
                    }
                }
            } else if (!P.close_encounter && P.inner_steps) {
                if (tests) {
                    check(S.ctx, red_check_collision(S.ctx, &handled));
                    if (handled) {
                        print_event_data(event_f, log_f, S);
                        check(S.ctx, red_clear_events(S.ctx));
                    }
                }
                int rebuilt = 0;
                check(S.ctx, red_check_rebuild(S.ctx, 8, &rebuilt));
                if (rebuilt) {
                    int m = 0;
                    check(S.ctx, red_last_rebuild_map(S.ctx, nullptr, 0, &m));
                    std::vector<int> map(std::max(m, 1));
                    check(S.ctx, red_last_rebuild_map(S.ctx, map.data(), m, &m));
                    std::vector<std::string> names(m);
                    std::vector<int> mt(m);
                    std::vector<double> ms(m);
                    for (int k = 0; k < m; k++) { names[k] = S.names[map[k]]; mt[k] = S.mig_type[map[k]]; ms[k] = S.mig_stop[map[k]]; }
                    S.names.swap(names); S.mig_type.swap(mt); S.mig_stop.swap(ms);
                    check(S.ctx, red_get_time(S.ctx, &t));
                    log_f << time_stamp() << " Rebuild the vectors and remove inactive bodies " << " t: " << std::scientific
                          << std::setprecision(16) << t << std::endl;
                }
            }
        };
        // fixed-step integrators run the iterations that need nothing from the host on the
        // device (red_run_fixed), in calls of at most 4096 steps so the info file keeps its
        // cadence; adaptive runs and unsupported configurations step one at a time
        bool fused = !opt.stepwise;
        while (t <= P.stop_time) {
            if (P.output_interval <= std::fabs(ps)) {
                ps = 0.0;
                print_result(result_f, S, P.binary);
            }
            if (fused) {
                int k = 0, why = 0;
                clock_t t0 = clock();
                const int rc = red_run_fixed(S.ctx, 4096, P.stop_time, P.output_interval, &ps, &k, &why);
                if (rc == RED_ERR_UNSUPPORTED) {
                    fused = false;
                } else {
                    check(S.ctx, rc);
                    if (k > 0) {
                        sum_steps += clock() - t0;
                        one_step = (clock() - t0) / k;
                        check(S.ctx, red_get_dt(S.ctx, &dt));
                    }
                    if (why == RED_FUSED_STOP_EVENTS) after_step();
                    // an output stop with no step done (output_interval <= 0) leaves the
                    // iteration to the host as an ejection stop does
                    if (why != RED_FUSED_STOP_EJECTION && !(why == RED_FUSED_STOP_OUTPUT && k == 0)) {
                        if (5.0 < (clock() - info_start) / (double)CLOCKS_PER_SEC)
                            print_info(info_f, S, dt, sum_steps, one_step, &info_start);
                        check(S.ctx, red_get_time(S.ctx, &t));
                        continue;
                    }
                }
            }
            int handled = 0;
            check(S.ctx, red_check_ejection_hit_centrum(S.ctx, &handled));
            if (handled) {
                print_event_data(event_f, log_f, S);
                check(S.ctx, red_clear_events(S.ctx));
            }
            clock_t t0 = clock();
            red_step_info info;
            check(S.ctx, red_step(S.ctx, &info));
            one_step = clock() - t0;
            sum_steps += one_step;
            dt = info.dt_did;
            ps += std::fabs(dt);
            after_step();
            if (5.0 < (clock() - info_start) / (double)CLOCKS_PER_SEC)
                print_info(info_f, S, dt, sum_steps, one_step, &info_start);
            check(S.ctx, red_get_time(S.ctx, &t));
        }
        print_info(info_f, S, dt, sum_steps, one_step, &info_start);
        if (0.0 < ps) print_result(result_f, S, P.binary);
    } catch (const Error& e) {
        if (log_f.is_open()) log_f << time_stamp() << "Error: " << e.msg << std::endl;
        std::cerr << "Error: " << e.msg << std::endl;
        exit_code = EXIT_FAILURE;
    }
    if (log_f.is_open()) log_f << time_stamp() << "Total time: " << (long)(time(nullptr) - start) << " s" << std::endl;
    std::cout << "Total time: " << (long)(time(nullptr) - start) << " s" << std::endl;
    if (S.ctx) red_ctx_destroy(S.ctx);
    return exit_code;
}
