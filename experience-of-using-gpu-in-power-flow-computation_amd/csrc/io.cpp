// CDF/CDFX reader and data.txt parser of libpf.
//
// Reader semantics follow CPU/leitura/cdf.cpp:3-493 and CPU/leitura/id2i.cpp:3-10:
//  * two header lines, bus records until a line containing "-999", one more
//    header line, branch records until the next "-999" (:182, :250);
//  * bus type 0/1 -> PQ, 2 -> PV, 3 -> slack; a second slack or any other
//    type is an error (:189-225);
//  * PV setpoint from the final-voltage field (:202); Q limits, loads and
//    generation divided by the integer baseMVA of the header (:175, :205-230);
//    the slack angle from degrees (:215); bus shunts taken as pu (:232-233);
//  * tap = a e^{j phi}, a = 0 meaning 1 (:273-280);
//  * a repeated (from, to) pair is merged into its first occurrence:
//    z = z1 z2 / (z1 + z2), bsh = 1 / (b1 + b2) (:289-313; the reciprocal is
//    the reference's and is kept for drop-in parity).
// CDFX moves every bus field after the 5-character id by one column and the
// branch fields after the two ids by two (:351-448).

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

#include "pf_internal.hpp"

namespace pf {

namespace {

constexpr double kPi = 3.14159265358979323846;

// std::string::substr(pos, n) semantics: pos beyond the end is an error.
bool field(const std::string &line, size_t pos, size_t n, std::string &out) {
    if (pos > line.size()) return false;
    out = line.substr(pos, n);
    return true;
}

bool fi(const std::string &line, size_t pos, size_t n, int &v) {
    std::string s;
    if (!field(line, pos, n, s)) return false;
    v = std::atoi(s.c_str());
    return true;
}

bool fd(const std::string &line, size_t pos, size_t n, double &v) {
    std::string s;
    if (!field(line, pos, n, s)) return false;
    v = std::atof(s.c_str());
    return true;
}

bool ends_section(const std::string &line) { return line.find("-999") != std::string::npos; }

void smith_div(double ar, double ai, double br, double bi, double &cr, double &ci) {
    // Smith, CACM Algorithm 116 (1962)
    if (std::fabs(br) >= std::fabs(bi)) {
        double r = bi / br, d = br + bi * r;
        cr = (ar + ai * r) / d;
        ci = (ai - ar * r) / d;
    } else {
        double r = br / bi, d = bi + br * r;
        cr = (ar * r + ai) / d;
        ci = (ai * r - ar) / d;
    }
}

}  // namespace

pf_status read_case(const std::string &path, double v0, double theta0, pf_network &net) {
    std::string ext;
    auto dot = path.find_last_of('.');
    if (dot != std::string::npos) ext = path.substr(dot + 1);
    for (auto &ch : ext) ch = (char)std::tolower((unsigned char)ch);
    bool wide;
    if (ext == "cdf") wide = false;
    else if (ext == "cdfx") wide = true;
    else if (ext == "m") return read_matpower(path, v0, theta0, net);     // CPU/leitura/arquivo.cpp:43-47
    else if (ext == "pwf") return read_pwf(path, v0, theta0, net);        // CPU/leitura/arquivo.cpp:37-41
    else return PF_INVALID_ARG;
    const size_t o = wide ? 1 : 0, ob = wide ? 2 : 0;

    std::ifstream f(path);
    if (!f) return PF_IO_ERROR;
    std::vector<std::string> bus_lines, br_lines;
    std::string line, header;
    if (!std::getline(f, header)) return PF_PARSE_ERROR;
    if (!std::getline(f, line)) return PF_PARSE_ERROR;
    bool ended = false;
    while (std::getline(f, line)) {
        if (ends_section(line)) { ended = true; break; }
        bus_lines.push_back(line);
    }
    if (!ended) return PF_PARSE_ERROR;
    std::getline(f, line);
    while (std::getline(f, line)) {
        if (ends_section(line)) break;
        br_lines.push_back(line);
    }

    net = pf_network();
    int mva = 0;
    if (!fi(header, 31, 6, mva) || mva == 0) return PF_PARSE_ERROR;
    net.baseMVA = mva;
    const int nB = (int)bus_lines.size();
    net.nB = nB;
    for (auto *a : {&net.V, &net.theta, &net.Pliq, &net.Qliq, &net.Pload, &net.Qload, &net.Pg,
                    &net.Qg, &net.Vbase, &net.gsh, &net.bsh})
        a->assign(nB, 0.0);
    net.id.assign(nB, 0);
    for (int i = 0; i < nB; i++) { net.V[i] = v0; net.theta[i] = theta0; }
    for (int i = 0; i < nB; i++) {
        const std::string &l = bus_lines[i];
        int type = 0;
        bool ok = fi(l, 0, 4 + o, net.id[i]) && fi(l, 24 + o, 2, type);
        if (!ok) return PF_PARSE_ERROR;
        if (type == 0 || type == 1) {
            net.pq.push_back(i);
        } else if (type == 2) {
            double v, qmax, qmin;
            ok = fd(l, 27 + o, 6, v) && fd(l, 90 + o, 7, qmax) && fd(l, 98 + o, 7, qmin);
            if (!ok) return PF_PARSE_ERROR;
            net.pv.push_back(i);
            net.V[i] = v;
            net.Vset.push_back(v);
            net.Qmax.push_back(qmax / net.baseMVA);
            net.Qmin.push_back(qmin / net.baseMVA);
        } else if (type == 3) {
            if (net.slack >= 0) return PF_PARSE_ERROR;   // "mais de uma barra swing"
            net.slack = i;
            double ang;
            ok = fd(l, 27 + o, 6, net.V[i]) && fd(l, 33 + o, 7, ang);
            if (!ok) return PF_PARSE_ERROR;
            net.theta[i] = ang * kPi / 180.;
        } else {
            return PF_PARSE_ERROR;                        // invalid bus type
        }
        ok = fd(l, 40 + o, 9, net.Pload[i]) && fd(l, 49 + o, 10, net.Qload[i]) &&
             fd(l, 59 + o, 8, net.Pg[i]) && fd(l, 67 + o, 8, net.Qg[i]) &&
             fd(l, 76 + o, 7, net.Vbase[i]) && fd(l, 114 + o, 8, net.bsh[i]) &&
             fd(l, 106 + o, 8, net.gsh[i]);
        if (!ok) return PF_PARSE_ERROR;
        net.Pload[i] /= net.baseMVA;
        net.Qload[i] /= net.baseMVA;
        net.Pg[i] /= net.baseMVA;
        net.Qg[i] /= net.baseMVA;
        net.Pliq[i] = net.Pg[i] - net.Pload[i];
        net.Qliq[i] = net.Qg[i] - net.Qload[i];
    }
    if (net.slack < 0) return PF_PARSE_ERROR;
    net.nPV = (int)net.pv.size();
    net.nPQ = (int)net.pq.size();

    std::unordered_map<int, int> id2i;
    id2i.reserve(nB * 2);
    for (int i = 0; i < nB; i++) id2i.emplace(net.id[i], i);   // first bus with an id wins
    std::unordered_map<long long, int> seen;
    seen.reserve(br_lines.size() * 2);
    for (const std::string &l : br_lines) {
        int a, b;
        double r, x, bc, tapA, ph;
        bool ok = fi(l, 0, 4 + o, a) && fi(l, 4 + ob, 5, b) && fd(l, 19 + ob, 10, r) &&
                  fd(l, 29 + ob, 11, x) && fd(l, 40 + ob, 10, bc) && fd(l, 76 + ob, 6, tapA) &&
                  fd(l, 83 + ob, 7, ph);
        if (!ok) return PF_PARSE_ERROR;
        auto ia = id2i.find(a), ib = id2i.find(b);
        if (ia == id2i.end() || ib == id2i.end()) return PF_PARSE_ERROR;   // "Barra inexistente!"
        int de = ia->second, pa = ib->second;
        long long key = ((long long)de << 32) | (unsigned)pa;
        auto it = seen.find(key);
        if (it == seen.end()) {
            seen.emplace(key, net.nL);
            net.from.push_back(de);
            net.to.push_back(pa);
            net.r.push_back(r);
            net.x.push_back(x);
            double phi = ph * kPi / 180.;
            net.phi.push_back(phi);
            double m = (tapA != 0.) ? tapA : 1.0;
            net.tap_re.push_back(m * std::cos(phi));
            net.tap_im.push_back(m * std::sin(phi));
            net.bch.push_back(bc);
            net.nL++;
        } else {
            int k = it->second;
            double nr = r * net.r[k] - x * net.x[k];
            double ni = r * net.x[k] + x * net.r[k];
            smith_div(nr, ni, net.r[k] + r, net.x[k] + x, net.r[k], net.x[k]);
            if (bc != 0 || net.bch[k] != 0) net.bch[k] = 1 / (bc + net.bch[k]);
            else net.bch[k] = 0;
            net.n_duplicates++;
        }
    }
    return PF_OK;
}

// ------------------------------------------------------------------ options

void options_default(pf_options &o) {
    // CPU/global.cpp:8-21, GPU/sistInfo.h:37-60
    std::memset(&o, 0, sizeof(o));
    o.metodo = PF_ESPARSO;
    o.lim_inj_reat = 0;
    o.openmp = 1;
    o.v_inicial = 1.0;
    o.theta_inicial = 0.0;
    o.tolerancia = 0.0001;
    o.no_max_iter = 100;
    std::strncpy(o.arq_entrada, "ieee14.cdf", sizeof(o.arq_entrada) - 1);
    o.verbose_mode = 0;
    o.laconic_mode = 0;
    o.output_bench = 0;
    o.output_ans = 1;
    o.output_iter = 0;
    o.streams = 1;
    o.precision = PF_FP64;
    o.refine_steps = 2;
    o.use_graph = 0;
    o.lu_blocked = 1;
}

namespace {

// Value lookup of CPU/leitura/config.cpp:21-23: the value starts `skip`
// characters after the first '=' following the key and ends at the next ';'.
// A missing key makes the search start near the top of the file (size_t
// wrap-around of npos + key length), so the first '=' of the file is used.
bool value_of(const std::string &s, const std::string &key, size_t skip, std::string &out) {
    size_t k = s.find(key);
    size_t from = k + key.size();               // wraps when k == npos, as in the reference
    size_t eq = s.find('=', from);
    size_t start = eq + skip;                   // wraps when eq == npos
    int inicio = (int)start;
    if (inicio < 0 || (size_t)inicio > s.size()) return false;
    size_t fim = s.find(';', (size_t)inicio);
    out = s.substr((size_t)inicio, fim - (size_t)inicio);
    return true;
}

}  // namespace

pf_status options_parse(const std::string &s, int cpu_layout, pf_options &o) {
    options_default(o);
    std::string v;
    if (!value_of(s, "laconic_mode", 2, v)) return PF_PARSE_ERROR;
    o.laconic_mode = (v == "OFF") ? 0 : 1;
    if (!value_of(s, "metodo", 2, v)) return PF_PARSE_ERROR;
    if (v == "esparso") o.metodo = PF_ESPARSO;
    else if (v == "denso") o.metodo = PF_DENSO;
    else if (v == "hibridoC") o.metodo = PF_HIBRIDO_C;
    else if (v == "hibridoA") o.metodo = PF_HIBRIDO_A;   // GPU/sistInfo.h:103-110
    else if (v == "hibridoB") o.metodo = PF_HIBRIDO_B;
    else return PF_INVALID_ARG;    // esparsoSimples, denso_LAPACKE: not provided
    if (!value_of(s, "lim_inj_reat", 2, v)) return PF_PARSE_ERROR;
    if (v == "OFF") o.lim_inj_reat = 0;
    else if (v == "ON") o.lim_inj_reat = 1;
    if (!value_of(s, "openmp", 2, v)) return PF_PARSE_ERROR;
    o.openmp = (v == "OFF") ? 0 : 1;
    if (!value_of(s, "v_inicial", 1, v)) return PF_PARSE_ERROR;
    o.v_inicial = std::atof(v.c_str());
    if (!value_of(s, "theta_inicial", 1, v)) return PF_PARSE_ERROR;
    o.theta_inicial = std::atof(v.c_str());
    if (!value_of(s, "tolerancia", 1, v)) return PF_PARSE_ERROR;
    o.tolerancia = std::atof(v.c_str());
    if (!value_of(s, "no_max_iter", 1, v)) return PF_PARSE_ERROR;
    try {
        o.no_max_iter = std::stoi(v);
    } catch (...) {
        return PF_PARSE_ERROR;
    }
    if (!value_of(s, "arq_entrada", 2, v)) return PF_PARSE_ERROR;
    std::string path = cpu_layout ? std::string("sistemas/") + v : v;
    std::strncpy(o.arq_entrada, path.c_str(), sizeof(o.arq_entrada) - 1);
    if (!value_of(s, "verbose_mode", 2, v)) return PF_PARSE_ERROR;
    if (v == "OFF") o.verbose_mode = 0;
    else if (v == "ON") o.verbose_mode = 1;
    if (!value_of(s, "output_bench", 2, v)) return PF_PARSE_ERROR;
    o.output_bench = (v == "ALL") ? 0 : (v == "FILE") ? 2 : 1;
    if (!value_of(s, "output_ans", 2, v)) return PF_PARSE_ERROR;
    o.output_ans = (v == "OFF") ? 0 : 1;
    if (!value_of(s, "output_iter", 2, v)) return PF_PARSE_ERROR;
    if (v == "OFF") o.output_iter = 0;
    else if (v == "ON") o.output_iter = 1;
    if (!value_of(s, "streams", 2, v)) return PF_PARSE_ERROR;
    if (v == "OFF") o.streams = 0;
    else if (v == "ON") o.streams = 1;
    // keys of this library, read only when present
    if (s.find("precisao") != std::string::npos && value_of(s, "precisao", 2, v)) {
        if (v == "FP64") o.precision = PF_FP64;
        else if (v == "MISTA" || v == "MIXED") o.precision = PF_MIXED;
        else return PF_PARSE_ERROR;
    }
    if (s.find("refinamentos") != std::string::npos && value_of(s, "refinamentos", 1, v))
        o.refine_steps = std::atoi(v.c_str());
    if (s.find("grafo") != std::string::npos && value_of(s, "grafo", 2, v))
        o.use_graph = (v == "ON") ? 1 : 0;
    if (s.find("lu_blocos") != std::string::npos && value_of(s, "lu_blocos", 2, v))
        o.lu_blocked = (v == "OFF") ? 0 : 1;
    return PF_OK;
}

}  // namespace pf
