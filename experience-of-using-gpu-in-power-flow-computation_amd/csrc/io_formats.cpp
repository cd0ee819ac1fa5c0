// MATPOWER (.m) and ANAREDE (.pwf) readers of libpf.
//
// MATPOWER, following lerArquivoMatPOWEReficiente + storeMatPOWER
// (CPU/leitura/m.cpp:246-536):
//  * line by line; '%' starts a comment, a trailing ';' is dropped, empty
//    lines are skipped (cleanLine, :226-230);
//  * "mpc.bus = [", "mpc.gen = [", "mpc.branch = [" open a matrix that ends
//    at the line holding ']'; rows are split at TABs (tokenize(line, "\t"));
//    "inf" reads as 999999 and "nan" as 0 (parseVector, :100-108);
//    "mpc.baseMVA = x" gives the base (quotes removed);
//  * bus type 1 -> PQ, 2 -> PV, 3 -> slack (a second one is an error), any
//    other type is an error; a PV or slack bus takes Pg, Qg (and a PV bus its
//    Q limits) from the FIRST generator row at its number; a PQ bus takes no
//    generation; |V| of PV and slack buses from VM, the slack angle from VA
//    in degrees; loads, generation, limits and the bus shunts GS/BS divided
//    by baseMVA (:403-455);
//  * every branch row is a branch (no parallel merge, status ignored):
//    z = R + jX, tap = (TAP, or 1 when 0) e^{j SHIFT}, line charging BR_B
//    (:457-520).
//  Deviations: a row shorter than the columns read (bus 13, gen 5, branch 10)
//  or a number that does not parse is PF_PARSE_ERROR (the reference indexes
//  past the row or throws); generator rows need not carry all 21 columns.
//
// ANAREDE, following getPWF / readPWF (CPU/leitura/pwf.cpp:3-516):
//  * sections DBAR, DLIN, DCTE, DGBT, DCSC start at a line holding only their
//    name and end at a line starting with "99999"; lines starting with '('
//    are comments (parse(), :386-478);
//  * fixed columns, blank fields taking the reference's defaults, numbers
//    without a '.' read with an implied decimal point (parseFloatPtoDecImp,
//    :200-208) and held as 32-bit floats as the reference's record structs
//    do (CPU/leitura/pwf.h:30-60);
//  * bus type '0' or '3' -> PQ, '1' -> PV (|V| and Q limits from the
//    record), '2' -> slack (|V| and angle), others an error; bus shunts are
//    0; base voltage from DGBT by the record's group field (:10-59);
//  * DLIN: z = (R + jX) / 100, tap = a e^{j phase}, line charging B / base;
//    DCSC series capacitors: z = j X / 100, tap 1, no charging; a repeated
//    (from, to) pair merges as in the CDF reader, its charging term being
//    1 / (B_raw + b_merged) with the raw susceptance (:61-197);
//  * the base from "BASE " in DCTE (default 100).
//  Deviations: a missing section is empty (the reference's search wraps), a
//  section ends before its "99999" line even when empty, CRLF line ends are
//  accepted, and a field past the end of a short line is blank.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <unordered_map>

#include "pf_internal.hpp"

namespace pf {

namespace {

constexpr double kPi = 3.14159265358979323846;

// ------------------------------------------------------------ MATPOWER

void clean_line(std::string &l) {
    const size_t pc = l.find('%');
    if (pc != std::string::npos) l.erase(pc);
    if (!l.empty() && l.back() == '\r') l.pop_back();
    // trailing blanks, at most one ';', then blanks again ("[ \t]*;?[ \t]*$")
    auto blank = [](char c) { return c == ' ' || c == '\t'; };
    while (!l.empty() && blank(l.back())) l.pop_back();
    if (!l.empty() && l.back() == ';') l.pop_back();
    while (!l.empty() && blank(l.back())) l.pop_back();
}

// "mpc.<name>" then blanks, '=', blanks and (for a matrix) '[': the position
// right after it, or npos
size_t after_assign(const std::string &l, const std::string &name, bool matrix) {
    const std::string key = "mpc." + name;
    for (size_t p = l.find(key); p != std::string::npos; p = l.find(key, p + 1)) {
        size_t q = p + key.size();
        while (q < l.size() && (l[q] == ' ' || l[q] == '\t')) q++;
        if (q >= l.size() || l[q] != '=') continue;
        q++;
        while (q < l.size() && (l[q] == ' ' || l[q] == '\t')) q++;
        if (!matrix) return q;
        if (q < l.size() && l[q] == '[') return q + 1;
    }
    return std::string::npos;
}

std::vector<std::string> split_tabs(const std::string &l) {
    std::vector<std::string> out;
    size_t p = 0;
    while (p < l.size()) {
        size_t e = l.find('\t', p);
        if (e == std::string::npos) e = l.size();
        if (e > p) out.push_back(l.substr(p, e - p));
        p = e + 1;
    }
    return out;
}

bool parse_number(std::string s, double &v) {
    auto replace_ci = [&](const char *pat, const char *with) {
        std::string low = s;
        for (auto &c : low) c = (char)std::tolower((unsigned char)c);
        for (size_t p = low.find(pat); p != std::string::npos; p = low.find(pat, p + std::strlen(with))) {
            s.replace(p, std::strlen(pat), with);
            low.replace(p, std::strlen(pat), with);
        }
    };
    replace_ci("inf", "999999");
    replace_ci("nan", "0");
    const char *b = s.c_str();
    char *e = nullptr;
    v = std::strtod(b, &e);
    return e != b;
}

// rows of a matrix that starts on `line` (after its '['), up to the line
// holding ']'
bool read_matrix(std::istream &f, std::string line, std::vector<std::vector<double>> &rows) {
    for (;;) {
        const bool last = line.find(']') != std::string::npos;
        if (last) line.erase(std::remove(line.begin(), line.end(), ']'), line.end());
        clean_line(line);
        if (!line.empty()) {
            std::vector<double> r;
            for (const std::string &t : split_tabs(line)) {
                double v;
                if (!parse_number(t, v)) return false;
                r.push_back(v);
            }
            rows.push_back(std::move(r));
        }
        if (last) return true;
        if (!std::getline(f, line)) return false;
    }
}

// ------------------------------------------------------------ ANAREDE

// std::string::substr with blanks past the end of the line
std::string col(const std::string &l, size_t pos, size_t n) {
    if (pos >= l.size()) return std::string(n, ' ');
    std::string s = l.substr(pos, n);
    if (s.size() < n) s.append(n - s.size(), ' ');
    return s;
}

bool is_blank(const std::string &s) {
    return std::all_of(s.begin(), s.end(), [](char c) { return std::isspace((unsigned char)c) != 0; });
}

std::string or_default(const std::string &s, const char *def) { return is_blank(s) ? std::string(def) : s; }

char char_or(const std::string &l, size_t pos, char def) {
    const char c = pos < l.size() ? l[pos] : '\0';
    return (c == '\0' || c == ' ') ? def : c;
}

// a number without '.' has `implied` decimal digits; the reference keeps it
// in a float
float implied_point(const std::string &s, int implied) {
    if (s.find('.') != std::string::npos) return (float)std::atof(s.c_str());
    return (float)(std::atoi(s.c_str()) / std::pow(10.0, implied));
}

// the records of a section: lines after the line "NAME" up to the line
// starting with "99999", comment lines ('(') skipped
std::vector<std::string> section(const std::vector<std::string> &lines, const std::string &name) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < lines.size() && lines[i] != name) i++;
    for (i++; i < lines.size(); i++) {
        const std::string &l = lines[i];
        if (l.compare(0, 5, "99999") == 0) break;
        if (!l.empty() && l[0] == '(') continue;
        out.push_back(l);
    }
    return out;
}

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

void init_buses(pf_network &net, int nB, double v0, double theta0) {
    net.nB = nB;
    for (auto *a : {&net.V, &net.theta, &net.Pliq, &net.Qliq, &net.Pload, &net.Qload, &net.Pg, &net.Qg,
                    &net.Vbase, &net.gsh, &net.bsh})
        a->assign(nB, 0.0);
    net.id.assign(nB, 0);
    std::fill(net.V.begin(), net.V.end(), v0);
    std::fill(net.theta.begin(), net.theta.end(), theta0);
}

}  // namespace

pf_status read_matpower(const std::string &path, double v0, double theta0, pf_network &net) {
    std::ifstream f(path);
    if (!f) return PF_IO_ERROR;
    std::vector<std::vector<double>> bus, gen, branch;
    std::string base_s;
    std::string line;
    while (std::getline(f, line)) {
        clean_line(line);
        if (line.empty()) continue;
        bool matrix = false;
        static const char *names[] = {"bus", "gen", "branch"};
        std::vector<std::vector<double>> *dst[] = {&bus, &gen, &branch};
        for (int k = 0; k < 3 && !matrix; k++) {
            const size_t p = after_assign(line, names[k], true);
            if (p == std::string::npos) continue;
            if (!read_matrix(f, line.substr(p), *dst[k])) return PF_PARSE_ERROR;
            matrix = true;
        }
        if (matrix) continue;
        const size_t p = after_assign(line, "baseMVA", false);
        if (p != std::string::npos) {
            base_s = line.substr(p);
            base_s.erase(std::remove_if(base_s.begin(), base_s.end(), [](char c) { return c == '"' || c == '\''; }),
                         base_s.end());
        }
    }
    double base = 0;
    if (base_s.empty() || !parse_number(base_s, base) || base == 0) return PF_PARSE_ERROR;
    for (auto &r : bus)
        if (r.size() < 13) return PF_PARSE_ERROR;
    for (auto &r : gen)
        if (r.size() < 5) return PF_PARSE_ERROR;
    for (auto &r : branch)
        if (r.size() < 10) return PF_PARSE_ERROR;

    net = pf_network();
    net.baseMVA = base;
    const int nB = (int)bus.size();
    init_buses(net, nB, v0, theta0);
    auto first_gen = [&](int id) {
        for (size_t g = 0; g < gen.size(); g++)
            if ((int)gen[g][0] == id) return (int)g;
        return -1;
    };
    for (int i = 0; i < nB; i++) {
        const auto &b = bus[i];
        net.id[i] = (int)b[0];
        int ig = -1;
        switch ((int)b[1]) {
        case 1:
            net.pq.push_back(i);
            break;
        case 2:
            net.pv.push_back(i);
            net.V[i] = b[7];
            net.Vset.push_back(b[7]);
            ig = first_gen(net.id[i]);
            net.Qmax.push_back(ig >= 0 ? gen[ig][3] / base : 0.0);
            net.Qmin.push_back(ig >= 0 ? gen[ig][4] / base : 0.0);
            break;
        case 3:
            if (net.slack >= 0) return PF_PARSE_ERROR;   // "mais de uma barra swing"
            net.slack = i;
            net.V[i] = b[7];
            net.theta[i] = b[8] * kPi / 180.;
            ig = first_gen(net.id[i]);
            break;
        default:
            return PF_PARSE_ERROR;                        // "valor de tipo de barra invalido"
        }
        net.Pload[i] = b[2] / base;
        net.Qload[i] = b[3] / base;
        net.Pg[i] = ig >= 0 ? gen[ig][1] / base : 0.0;
        net.Qg[i] = ig >= 0 ? gen[ig][2] / base : 0.0;
        net.Vbase[i] = b[9];
        net.bsh[i] = b[5] / base;
        net.gsh[i] = b[4] / base;
        net.Pliq[i] = net.Pg[i] - net.Pload[i];
        net.Qliq[i] = net.Qg[i] - net.Qload[i];
    }
    if (net.slack < 0) return PF_PARSE_ERROR;
    net.nPV = (int)net.pv.size();
    net.nPQ = (int)net.pq.size();
    std::unordered_map<int, int> id2i;
    for (int i = 0; i < nB; i++) id2i.emplace(net.id[i], i);   // first bus with a number wins
    for (const auto &r : branch) {
        auto ia = id2i.find((int)r[0]), ib = id2i.find((int)r[1]);
        if (ia == id2i.end() || ib == id2i.end()) return PF_PARSE_ERROR;
        net.from.push_back(ia->second);
        net.to.push_back(ib->second);
        net.r.push_back(r[2]);
        net.x.push_back(r[3]);
        net.bch.push_back(r[4]);
        const double phi = r[9] * kPi / 180., a = r[8] != 0. ? r[8] : 1.0;
        net.phi.push_back(phi);
        net.tap_re.push_back(a * std::cos(phi));
        net.tap_im.push_back(a * std::sin(phi));
        net.nL++;
    }
    return PF_OK;
}

pf_status read_pwf(const std::string &path, double v0, double theta0, pf_network &net) {
    std::ifstream f(path);
    if (!f) return PF_IO_ERROR;
    std::vector<std::string> lines;
    for (std::string l; std::getline(f, l);) {
        if (!l.empty() && l.back() == '\r') l.pop_back();
        lines.push_back(l);
    }
    const auto dbar = section(lines, "DBAR"), dlin = section(lines, "DLIN"), dcte = section(lines, "DCTE"),
               dgbt = section(lines, "DGBT"), dcsc = section(lines, "DCSC");
    float base = 100.0f;
    for (const std::string &l : dcte) {
        const size_t p = l.find("BASE ");
        if (p != std::string::npos) {
            base = (float)std::atof(col(l, p + 5, 6).c_str());
            break;
        }
    }
    std::map<std::string, float> kv;   // DGBT: group -> base kV
    for (const std::string &l : dgbt) {
        std::string key = col(l, 0, 2);
        key.erase(std::remove_if(key.begin(), key.end(), [](char c) { return std::isspace((unsigned char)c); }),
                  key.end());
        kv[key] = (float)std::atof(col(l, 3, 7).c_str());
    }
    net = pf_network();
    net.baseMVA = base;
    const double B = base;   // sistema.baseMVA: a double in the reference
    const int nB = (int)dbar.size();
    init_buses(net, nB, v0, theta0);
    for (int i = 0; i < nB; i++) {
        const std::string &l = dbar[i];
        net.id[i] = std::atoi(col(l, 0, 5).c_str());
        const char type = char_or(l, 7, '0');
        const std::string group = or_default(col(l, 8, 2), "0");
        const float v = implied_point(or_default(col(l, 24, 4), "1000"), 3);
        const float ang = (float)std::atof(or_default(col(l, 28, 4), "0.0").c_str());
        const float pg = (float)std::atof(or_default(col(l, 32, 5), "0.0").c_str());
        const float qg = (float)std::atof(or_default(col(l, 37, 5), "0.0").c_str());
        const float qmin = (float)std::atof(or_default(col(l, 42, 5), "-9999.0").c_str());
        const float qmax = (float)std::atof(or_default(col(l, 47, 5), "99999.0").c_str());
        const float pl = (float)std::atof(or_default(col(l, 58, 5), "0.0").c_str());
        const float ql = (float)std::atof(or_default(col(l, 63, 5), "0.0").c_str());
        switch (type) {
        case '3':   // bus with voltage limits: a PQ bus here
        case '0':
            net.pq.push_back(i);
            break;
        case '1':
            net.pv.push_back(i);
            net.V[i] = v;
            net.Vset.push_back(v);
            net.Qmax.push_back(qmax / B);
            net.Qmin.push_back(qmin / B);
            break;
        case '2':
            if (net.slack >= 0) return PF_PARSE_ERROR;   // "mais de uma barra swing"
            net.slack = i;
            net.V[i] = v;
            net.theta[i] = ang * kPi / 180.;
            break;
        default:
            return PF_PARSE_ERROR;
        }
        net.Pload[i] = pl / B;
        net.Qload[i] = ql / B;
        net.Pg[i] = pg / B;
        net.Qg[i] = qg / B;
        auto it = kv.find(group);   // the group as written (the DGBT keys are trimmed)
        net.Vbase[i] = it != kv.end() ? it->second : 0.0f;
        net.Pliq[i] = net.Pg[i] - net.Pload[i];
        net.Qliq[i] = net.Qg[i] - net.Qload[i];
    }
    if (net.slack < 0) return PF_PARSE_ERROR;
    net.nPV = (int)net.pv.size();
    net.nPQ = (int)net.pq.size();
    std::unordered_map<int, int> id2i;
    for (int i = 0; i < nB; i++) id2i.emplace(net.id[i], i);
    std::unordered_map<long long, int> seen;
    // one branch: z, tap, line charging (already per unit) and the raw
    // susceptance the merge rule uses
    auto add = [&](int a, int b, double r, double x, double tap, double phase_deg, double bsh, double b_raw) {
        auto ia = id2i.find(a), ib = id2i.find(b);
        if (ia == id2i.end() || ib == id2i.end()) return false;
        const int de = ia->second, pa = ib->second;
        const long long key = ((long long)de << 32) | (unsigned)pa;
        auto it = seen.find(key);
        if (it == seen.end()) {
            seen.emplace(key, net.nL);
            net.from.push_back(de);
            net.to.push_back(pa);
            net.r.push_back(r);
            net.x.push_back(x);
            const double phi = phase_deg * kPi / 180.;
            net.phi.push_back(phi);
            net.tap_re.push_back(tap * std::cos(phi));
            net.tap_im.push_back(tap * std::sin(phi));
            net.bch.push_back(bsh);
            net.nL++;
        } else {
            const int k = it->second;
            const double nr = r * net.r[k] - x * net.x[k], ni = r * net.x[k] + x * net.r[k];
            smith_div(nr, ni, net.r[k] + r, net.x[k] + x, net.r[k], net.x[k]);
            if (b_raw != 0 || net.bch[k] != 0) net.bch[k] = 1 / (b_raw + net.bch[k]);
            else net.bch[k] = 0;
            net.n_duplicates++;
        }
        return true;
    };
    for (const std::string &l : dlin) {
        const int a = std::atoi(col(l, 0, 5).c_str()), b = std::atoi(col(l, 10, 5).c_str());
        const float r = implied_point(or_default(col(l, 20, 6), "0.0"), 2);
        const float x = implied_point(col(l, 26, 6), 2);
        const float bs = implied_point(or_default(col(l, 32, 6), "0.0"), 3);
        const float tap = implied_point(or_default(col(l, 38, 5), "1.0"), 3);
        const float ph = implied_point(col(l, 53, 5), 2);
        // float record fields over an int 100 divide in float (CPU/leitura/pwf.cpp:85-86)
        if (!add(a, b, r / 100, x / 100, tap, ph, bs / B, bs)) return PF_PARSE_ERROR;
    }
    for (const std::string &l : dcsc) {
        const int a = std::atoi(col(l, 0, 5).c_str()), b = std::atoi(col(l, 9, 5).c_str());
        const float xc = (float)std::atof(or_default(col(l, 45, 6), "0.0").c_str());
        if (!add(a, b, 0.0, xc / 100, 1.0, 0.0, 0.0, 0.0)) return PF_PARSE_ERROR;
    }
    return PF_OK;
}

}  // namespace pf
