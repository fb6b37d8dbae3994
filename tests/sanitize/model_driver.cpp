// Drives the product's equation compiler (gopf_amd/csrc/model.cpp) from a line protocol on
// stdin, for the sanitizer test (tests/test_sanitizers.py), which builds it once plainly and
// once with AddressSanitizer and UndefinedBehaviorSanitizer and compares the two outputs.
//   F <name>            Model.AddField (8 nodes)
//   S <name> <re> <im>  Model.AddScalar
//   P <name> <field> <coef> <power>   Model.RegisterFunction as a one-term polynomial
//   E <equation>        Model.AddEquation
//   I                   Model.Init, then print every equation's terms and denominators
//   R                   start a new model
// Every panic of the reference (a ParseError or a runtime error here) prints "PANIC <msg>"
// and starts a new model.
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>

#include "model.h"

using namespace gpf;

static void print_terms(const char* what, int eq, const std::vector<TermDesc>& ts) {
    for (const TermDesc& t : ts)
        std::printf("%s %d kind %d coef %.17g %.17g brick %d lap %d has_lap %d user %d sign %g\n", what, eq, t.kind,
                    t.coef.real(), t.coef.imag(), t.brick, t.lap, t.has_lap ? 1 : 0, t.user, t.sign);
}

int main() {
    Model m;
    std::string line;
    while (std::getline(std::cin, line)) {
        if (line.empty()) continue;
        const char op = line[0];
        const std::string rest = line.size() > 2 ? line.substr(2) : "";
        try {
            if (op == 'F') {
                m.add_field(rest, std::vector<std::complex<double>>(8, {0.0, 0.0}));
            } else if (op == 'S') {
                std::istringstream is(rest);
                std::string name;
                double re = 0, im = 0;
                is >> name >> re >> im;
                m.add_scalar(name, {re, im});
            } else if (op == 'P') {
                std::istringstream is(rest);
                std::string name, field;
                double coef = 0, power = 0;
                is >> name >> field >> coef >> power;
                std::vector<double> pw(m.field_names.size(), 0.0);
                for (size_t f = 0; f < m.field_names.size(); ++f)
                    if (m.field_names[f] == field) pw[f] = power;
                m.register_polynomial(name, {{coef, pw}});
            } else if (op == 'E') {
                m.add_equation(rest);
            } else if (op == 'I') {
                m.init();
                for (size_t e = 0; e < m.rhs.size(); ++e) {
                    print_terms("term", (int)e, m.rhs[e].first);
                    print_terms("denum", (int)e, m.rhs[e].second);
                }
                for (const std::string& n : m.all_field_names()) std::printf("brick %s\n", n.c_str());
            } else if (op == 'R') {
                m = Model();
            }
        } catch (const std::exception& ex) {
            std::printf("PANIC %s\n", ex.what());
            m = Model();
        }
    }
    return 0;
}
