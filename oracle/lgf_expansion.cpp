// Oracle evaluation of the far-field LGF / Mehrstellen expansions (reference:
// src/lgf_3unb0spe.hpp:5-374), read from the coefficient table
// flups_amd/kernel/lgf_expansion.txt: each line "name a b c p l coef" adds
// coef * n1^a n2^b n3^c n^-p log(n)^l. Test infrastructure only.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <vector>

#include "oracle.hpp"

namespace oracle {

namespace {
struct Term { int a, b, c, p, l; double coef; };
std::mutex g_mtx;
std::map<std::string, std::vector<Term>> g_terms;

const std::vector<Term>& terms(const std::string& kdir, const std::string& name) {
    std::lock_guard<std::mutex> lk(g_mtx);
    if (g_terms.empty()) {
        std::ifstream f(kdir + "/lgf_expansion.txt");
        if (!f) throw std::runtime_error("unable to read lgf_expansion.txt in " + kdir);
        std::string line;
        while (std::getline(f, line)) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream is(line);
            std::string nm;
            Term t;
            is >> nm >> t.a >> t.b >> t.c >> t.p >> t.l >> t.coef;
            g_terms[nm].push_back(t);
        }
    }
    auto it = g_terms.find(name);
    if (it == g_terms.end()) throw std::runtime_error("no expansion " + name);
    return it->second;
}

double eval(const std::vector<Term>& ts, double n1, double n2, double n3) {
    const double n = std::sqrt(n1 * n1 + n2 * n2 + n3 * n3);
    const double ln = std::log(n);
    double s = 0.0;
    for (const Term& t : ts)
        s += t.coef * std::pow(n1, t.a) * std::pow(n2, t.b) * std::pow(n3, t.c) * std::pow(n, -t.p) * (t.l ? ln : 1.0);
    return s;
}
}  // namespace

double kernel_expansion(const std::string& kdir, const std::string& name, double n1, double n2, double n3) {
    return eval(terms(kdir, name), n1, n2, n3);
}
double lgf3_expansion(const std::string& kdir, int order, double n1, double n2, double n3) {
    return eval(terms(kdir, "lgf_" + std::to_string(order) + "_3unb0spe"), n1, n2, n3);
}
double lgf2_expansion(const std::string& kdir, int order, double n1, double n2) {
    return eval(terms(kdir, "lgf_" + std::to_string(order) + "_2unb0spe"), n1, n2, 0.0);
}

}  // namespace oracle
