// model.cpp -- gopf's equation parser, restated in C++.
//
// Reference: pf/model.go:140-260, pf/rhsBuilder.go:26-287, pf/util.go:17-213.
// Go's regexp is leftmost-first like ECMAScript std::regex for every pattern
// used here; Go's FindAll rule that drops an empty match abutting the previous
// match (regexp.go allMatches) is restated in go_find_all.
#include "model.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <regex>

namespace gpf {

namespace {

std::regex compile(const std::string& pat) {
    try {
        return std::regex(pat, std::regex::ECMAScript);
    } catch (const std::regex_error& e) {
        throw ParseError("regexp: Compile(" + pat + "): " + e.what());
    }
}

std::vector<std::string> go_find_all(const std::string& pat, const std::string& s) {
    const std::regex rx = compile(pat);
    std::vector<std::string> out;
    size_t pos = 0;
    long prev_end = -1;
    const size_t end = s.size();
    while (pos <= end) {
        std::smatch m;
        auto flags = std::regex_constants::match_default;
        if (pos > 0) flags |= std::regex_constants::match_prev_avail;
        if (!std::regex_search(s.begin() + pos, s.end(), m, rx, flags)) break;
        const size_t ms = pos + m.position(0), me = ms + m.length(0);
        bool accept = true;
        if (me == pos) {
            if ((long)ms == prev_end) accept = false;
            pos = (pos < end) ? pos + 1 : end + 1;
        } else {
            pos = me;
        }
        prev_end = (long)me;
        if (accept) out.push_back(m.str(0));
    }
    return out;
}

std::string go_find_string(const std::string& pat, const std::string& s) {
    const std::regex rx = compile(pat);
    std::smatch m;
    if (!std::regex_search(s, m, rx)) return "";
    return m.str(0);
}

std::vector<std::string> go_find_submatch(const std::string& pat, const std::string& s) {
    const std::regex rx = compile(pat);
    std::smatch m;
    std::vector<std::string> out;
    if (!std::regex_search(s, m, rx)) return out;
    for (size_t i = 0; i < m.size(); ++i) out.push_back(m[i].matched ? m.str(i) : std::string());
    return out;
}

bool parse_float(const std::string& s, double* v) {
    if (s.empty()) return false;
    char* endp = nullptr;
    *v = std::strtod(s.c_str(), &endp);
    return endp && *endp == '\0';
}

std::string replace_all(std::string s, const std::string& from, const std::string& to) {
    if (from.empty()) return s;
    size_t pos = 0;
    while ((pos = s.find(from, pos)) != std::string::npos) {
        s.replace(pos, from.size(), to);
        pos += to.size();
    }
    return s;
}

std::vector<std::string> split(const std::string& s, const std::string& sep) {
    std::vector<std::string> out;
    size_t start = 0;
    while (true) {
        size_t p = s.find(sep, start);
        if (p == std::string::npos) {
            out.push_back(s.substr(start));
            break;
        }
        out.push_back(s.substr(start, p - start));
        start = p + sep.size();
    }
    return out;
}

bool has_prefix(const std::string& s, const std::string& p) { return s.compare(0, p.size(), p) == 0; }

const std::vector<std::string>& known_prefixes() {
    static const std::vector<std::string> k = {"-", "LAP^4", "LAP^2", "LAP", "*"};
    return k;
}

void panic_on_prefix_in_name(const std::string& name) {
    if (!get_known_prefixes(name).empty())
        throw ParseError("The words [- LAP^4 LAP^2 LAP *] are reserved. Do not include them in your variable names");
}

std::string get_non_linear_field_expressions(const std::string& pattern, const std::string& field,
                                             const std::vector<std::string>& names) {
    std::string expr;
    for (const auto& n : names) {
        if (n == field && is_bilinear(pattern, field, names)) continue;
        std::string res = go_find_string(n + "[^\\*]*", pattern);
        if (!res.empty()) expr += res + "*";
    }
    if (expr.size() > 1) return expr.substr(0, expr.size() - 1);
    return expr;
}

}  // namespace

std::complex<double> go_cmul(std::complex<double> x, std::complex<double> y) {
    const double a = x.real(), b = x.imag(), c = y.real(), d = y.imag();
    return {a * c - b * d, a * d + b * c};
}

std::complex<double> go_cpow(std::complex<double> x, double p) {
    if (x.real() == 0.0 && x.imag() == 0.0) {
        if (p == 0.0) return {1.0, 0.0};
        if (p < 0.0) return {INFINITY, 0.0};
        return {0.0, 0.0};
    }
    const double mod = std::hypot(x.real(), x.imag());
    if (mod == 0.0) return {0.0, 0.0};
    const double r = std::pow(mod, p);
    const double th = p * std::atan2(x.imag(), x.real());
    return {r * std::cos(th), r * std::sin(th)};
}

// SplitOnMany (util.go:166-198): breadth-first queue of substrings
std::vector<SubStringDelimiter> split_on_many(const std::string& value, const std::vector<std::string>& delims) {
    std::vector<SubStringDelimiter> substrings;
    if (value.empty()) throw ParseError("runtime error: index out of range [0] with length 0");
    std::string first;
    for (const auto& d : delims)
        if (value.substr(0, 1) == d) { first = d; break; }
    std::vector<SubStringDelimiter> queue = {{value, first}};
    std::string all;
    for (const auto& d : delims) all += d;
    size_t head = 0;
    while (head < queue.size()) {
        SubStringDelimiter cur = queue[head++];
        if (cur.sub.find_first_of(all) == std::string::npos) {
            substrings.push_back(cur);
            continue;
        }
        std::string delim = delims[0];
        for (const auto& d : delims)
            if (cur.sub.find(d) != std::string::npos) { delim = d; break; }
        std::vector<std::string> parts;
        for (auto& s : split(cur.sub, delim))
            if (!s.empty()) parts.push_back(s);
        if (parts.empty()) throw ParseError("runtime error: index out of range [0] with length 0");
        queue.push_back({parts[0], cur.delim});
        for (size_t i = 1; i < parts.size(); ++i) queue.push_back({parts[i], delim});
    }
    return substrings;
}

std::string field_name_from_leibniz(const std::string& l) {
    if (l.size() <= 3) throw ParseError("rhsbuilder: Length of the Leibniz formatted string has to be at least 3");
    if (l.substr(0, 1) != "d" || l.substr(l.size() - 3) != "/dt")
        throw ParseError("rhsbuilder: Passed string is not a leibniz formatted string");
    return l.substr(1, l.size() - 4);
}

bool is_bilinear(const std::string& term, const std::string& field, const std::vector<std::string>& names) {
    if (go_find_all(field, term).size() != 1) return false;
    for (const auto& f : names) {
        if (f == field) continue;
        if (!go_find_all(f, term).empty()) return false;
    }
    const std::string res = go_find_string(field + "*[^/\\*]*", term);
    const auto m = go_find_submatch("\\^(-?\\d+\\.?\\d*)", res);
    if (m.size() <= 1) return true;
    double p;
    if (!parse_float(m[1], &p)) return true;
    return std::fabs(p - 1.0) < 1e-10;
}

std::string remove_known_prefixes(std::string s) {
    const auto& k = known_prefixes();
    size_t i = 0;
    while (i < k.size()) {
        if (has_prefix(s, k[i])) {
            s = s.substr(k[i].size());
            i = 0;
        } else {
            ++i;
        }
    }
    return s;
}

std::vector<std::string> get_known_prefixes(std::string s) {
    std::vector<std::string> pref;
    const auto& k = known_prefixes();
    size_t i = 0;
    while (i < k.size()) {
        if (has_prefix(s, k[i])) {
            pref.push_back(k[i]);
            s = s.substr(k[i].size());
            i = 0;
        } else {
            ++i;
        }
    }
    return pref;
}

std::string sort_factors(const std::string& expr) {
    auto parts = split(expr, "*");
    std::sort(parts.begin(), parts.end());
    std::string out;
    for (size_t i = 0; i < parts.size(); ++i) {
        if (i) out += "*";
        out += parts[i];
    }
    return out;
}

std::string get_field_name(const std::string& term, const std::vector<std::string>& names) {
    std::string field;
    for (const auto& f : names) {
        if (term.find(f) == std::string::npos) continue;
        const std::string without = replace_all(term, f, "");
        bool ok = true;
        for (const auto& f1 : names)
            if (without.find(f1) != std::string::npos) { ok = false; break; }
        if (ok && f.size() > field.size()) field = f;
    }
    return field;
}

double get_power(const std::string& pattern) {
    const auto m = go_find_submatch("\\^(-?\\d+\\.?\\d*)", pattern);
    if (m.size() <= 1) return 1.0;
    double p;
    if (!parse_float(m[1], &p)) throw ParseError("strconv.ParseFloat: parsing \"" + m[1] + "\": invalid syntax");
    return p;
}

double coef_derivative(const TermDesc& t, const std::map<std::string, std::complex<double>>& scalars,
                       const std::string& name) {
    double d = 0.0;
    for (size_t j = 0; j < t.factors.size(); ++j) {
        if (t.factors[j].first != name) continue;
        const double sv = scalars.at(name).real();
        const double p = t.factors[j].second;
        double v = t.sign * p * std::pow(sv, p - 1.0);
        for (size_t i = 0; i < t.factors.size(); ++i)
            if (i != j) v *= std::pow(scalars.at(t.factors[i].first).real(), t.factors[i].second);
        d += v;
    }
    return d;
}

// ---------------------------------------------------------------------------
// Model
// ---------------------------------------------------------------------------

std::vector<std::string> Model::all_field_names() const {
    std::vector<std::string> n = field_names;
    for (const auto& d : derived) n.push_back(d.name);
    return n;
}

bool Model::is_field_name(const std::string& n) const {
    for (const auto& f : field_names)
        if (f == n) return true;
    for (const auto& d : derived)
        if (d.name == n) return true;
    return false;
}

bool Model::is_brick_name(const std::string& n) const { return is_field_name(n) || scalars.count(n) > 0; }

int Model::brick_index(const std::string& n) const {
    auto names = all_field_names();
    for (size_t i = 0; i < names.size(); ++i)
        if (names[i] == n) return (int)i;
    throw ParseError("unknown brick " + n);
}

long Model::num_nodes() const {
    if (field_data.empty()) throw ParseError("Model: No fields added");
    return (long)field_data[0].size();
}

void Model::add_field(const std::string& name, std::vector<std::complex<double>> data) {
    panic_on_prefix_in_name(name);
    field_names.push_back(name);
    field_data.push_back(std::move(data));
}

void Model::add_scalar(const std::string& name, std::complex<double> v) {
    panic_on_prefix_in_name(name);
    scalars[name] = v;
}

void Model::register_explicit_term(const std::string& name, const UserTerm& t) {
    panic_on_prefix_in_name(name);
    explicit_terms[name] = t;
}

void Model::register_implicit_term(const std::string& name, const UserTerm& t) {
    panic_on_prefix_in_name(name);
    implicit_terms[name] = t;
}

void Model::register_polynomial(const std::string& name,
                                const std::vector<std::pair<double, std::vector<double>>>& terms) {
    panic_on_prefix_in_name(name);
    if (field_data.empty()) throw ParseError("runtime error: index out of range [0] with length 0");
    DerivedDesc d;
    d.name = name;
    d.kind = 1;
    for (const auto& t : terms) {
        std::vector<std::pair<int, double>> facs;
        for (size_t f = 0; f < t.second.size() && f < field_names.size(); ++f)
            if (t.second[f] != 0.0) facs.push_back({(int)f, t.second[f]});
        d.terms.push_back({t.first, facs});
    }
    derived.push_back(d);
}

void Model::register_white_noise(const std::string& name, double strength, unsigned long long seed) {
    panic_on_prefix_in_name(name);
    if (field_data.empty()) throw ParseError("runtime error: index out of range [0] with length 0");
    DerivedDesc d;
    d.name = name;
    d.kind = 2;
    d.noise_std = std::sqrt(2.0 * strength);
    d.noise_seed = seed;
    derived.push_back(d);
}

void Model::register_conservative_noise(const std::string& term, double strength, int dim, unsigned long long seed) {
    panic_on_prefix_in_name(term);
    if (dim < 1 || dim > 3) throw ParseError("ConservativeNoise: Dim must be 1, 2 or 3");
    if (field_data.empty()) throw ParseError("runtime error: index out of range [0] with length 0");
    UserTerm u;
    u.kind = kConservativeNoise;
    u.cn_dim = dim;
    // the reference prefixes the names with a random uint32 (noise.go:35-46); the seed's low
    // 32 bits play that part here
    const unsigned prefix = (unsigned)(seed & 0xffffffffu);
    for (int c = 0; c < dim; ++c) {
        DerivedDesc d;
        d.name = std::to_string(prefix) + "_current_" + std::to_string(c);
        d.kind = 2;
        d.noise_std = std::sqrt(2.0 * strength);
        d.noise_seed = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(c + 1);
        u.cn_names[c] = d.name;
        derived.push_back(d);
    }
    explicit_terms[term] = u;
}

void Model::add_equation(const std::string& eq0) {
    const std::string eq = replace_all(eq0, " ", "");
    equations.push_back(eq);
    update_derived_fields(eq);
}

// UpdateDerivedFields (model.go:165-195)
void Model::update_derived_fields(const std::string& eq) {
    auto sides = split(eq, "=");
    if (sides.size() < 2) throw ParseError("runtime error: index out of range [1] with length 1");
    std::vector<std::string> splitted;
    for (const auto& s : split_on_many(sides[1], {"+", "-"})) splitted.push_back(s.sub);
    const std::string field = field_name_from_leibniz(sides[0]);
    const std::vector<std::string> names = field_names;
    for (const auto& s : splitted) {
        if (implicit_terms.count(s) || explicit_terms.count(s)) continue;
        const std::string nf = sort_factors(get_non_linear_field_expressions(s, field, names));
        if (!nf.empty() && !is_field_name(nf)) {
            if (field_data.empty()) throw ParseError("runtime error: index out of range [0] with length 0");
            derived.push_back(derived_from_desc(nf));
        }
    }
}

// DerivedFieldCalcFromDesc (util.go:38-65)
DerivedDesc Model::derived_from_desc(const std::string& desc) const {
    DerivedDesc d;
    d.name = desc;
    d.kind = 0;
    std::vector<std::pair<int, double>> facs;
    for (const auto& part : go_find_all("[^\\*]*", desc)) {
        const std::string name = go_find_string("^[^\\^]*", part);
        int idx = -1;
        for (size_t f = 0; f < field_names.size(); ++f)
            if (field_names[f] == name) idx = (int)f;
        if (idx < 0) throw ParseError("derived field " + desc + ": " + name + " is not a field");
        facs.push_back({idx, get_power(part)});
    }
    d.terms.push_back({1.0, facs});
    return d;
}

// ValidName (rhsBuilder.go:108-122)
bool Model::valid_name(const std::string& name) const {
    std::string s = replace_all(name, " ", "");
    if (s.empty() || s == "LAP") return true;
    if (s.size() < 3) throw ParseError("runtime error: slice bounds out of range [:3] with length " + std::to_string(s.size()));
    if (s.substr(0, 3) == "LAP") s = s.substr(3);
    return is_brick_name(s) || is_field_name(s);
}

// ConcreteTerm (rhsBuilder.go:125-190)
TermDesc Model::concrete_term(const SubStringDelimiter& td) {
    const std::string& term = td.sub;
    const double sign = (td.delim == "-") ? -1.0 : 1.0;
    std::vector<std::string> bricks;
    std::vector<double> powers;
    for (const auto& part : go_find_all("[^\\*]*", term)) {
        const std::string name = go_find_string("^[^\\^]*", part);
        if (!is_field_name(name) && is_brick_name(name)) {
            bricks.push_back(name);
            powers.push_back(get_power(part));
        } else if (!valid_name(name)) {
            throw ParseError("rhsBuilder: Name " + name + " is not defined!");
        }
    }
    const std::string fname = get_field_name(sort_factors(term), all_field_names());
    TermDesc t;
    t.kind = kConcrete;
    std::complex<double> coef(sign, 0.0);
    for (size_t j = 0; j < bricks.size(); ++j) coef = go_cmul(coef, go_cpow(scalars.at(bricks[j]), powers[j]));
    t.coef = coef;
    t.sign = sign;
    for (size_t j = 0; j < bricks.size(); ++j) t.factors.push_back({bricks[j], powers[j]});
    if (!fname.empty()) t.brick = brick_index(fname);
    if (term.find("LAP") != std::string::npos) {
        const std::string res = go_find_string("LAP*[^a-zA-Z]*", term);
        t.has_lap = true;
        t.lap = (int)get_power(res);
    }
    return t;
}

// constructFunc (rhsBuilder.go:199-242): "-" flips the sign, LAP^n multiplies by L^n
TermDesc Model::user_term(const UserTerm& u, const std::vector<std::string>& prefixes) {
    double sign = 1.0;
    int lap = 0;
    for (const auto& p : prefixes) {
        if (p == "-") sign = -sign;
        else if (p == "LAP^4") lap += 4;
        else if (p == "LAP^2") lap += 2;
        else if (p == "LAP") lap += 1;
    }
    if (u.kind == kVolumeConserving && !is_field_name(u.indicator) &&
        std::none_of(derived.begin(), derived.end(), [&](const DerivedDesc& d) { return d.name == u.indicator; }))
        throw ParseError("VolumeConservingLP: Indicator is not a derived field");
    auto& table = user_table[u.kind];
    table.push_back(u);
    TermDesc t;
    t.kind = u.kind;
    t.coef = {sign, 0.0};
    t.lap = lap;
    t.user = (int)table.size() - 1;
    return t;
}

// Build (rhsBuilder.go:26-54)
std::pair<std::vector<TermDesc>, std::vector<TermDesc>> Model::build(const std::string& eq) {
    auto sides = split(eq, "=");
    if (sides.size() != 2) throw ParseError("build: equality sign can only occur once");
    const std::string field = field_name_from_leibniz(sides[0]);
    std::vector<TermDesc> terms, denum;
    for (const auto& t : split_on_many(sides[1], {"+", "-"})) {
        const std::string name = remove_known_prefixes(t.sub);
        auto prefixes = get_known_prefixes(t.sub);
        prefixes.push_back(t.delim);
        if (implicit_terms.count(name)) {
            denum.push_back(user_term(implicit_terms.at(name), prefixes));
        } else if (explicit_terms.count(name)) {
            terms.push_back(user_term(explicit_terms.at(name), prefixes));
        } else if (is_bilinear(t.sub, field, all_field_names())) {
            denum.push_back(concrete_term({replace_all(t.sub, field, ""), t.delim}));
        } else {
            terms.push_back(concrete_term(t));
        }
    }
    return {terms, denum};
}

// Model.Init (model.go:244-260)
void Model::init() {
    rhs.clear();
    for (auto& t : user_table) t.clear();
    for (const auto& eq : equations) rhs.push_back(build(eq));
}

}  // namespace gpf
