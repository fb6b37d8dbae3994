// model.h -- host mirror of gopf's pf.Model and its equation parser.
//
// The classification of equation terms into explicit (RHS.Terms) and
// implicit (RHS.Denum) parts follows pf/rhsBuilder.go:26-242 and
// pf/util.go:17-104 statement by statement; each closure of the reference
// becomes a TermDesc record the device plan compiles into kernel arguments.
#pragma once

#include <complex>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace gpf {

// Raised where the reference panics (panic / log.Fatalf / Go runtime error).
struct ParseError : std::runtime_error {
    explicit ParseError(const std::string& m) : std::runtime_error(m) {}
};

enum TermKind {
    kConcrete = 0,
    kSquaredGradient = 1,
    kHomogeneousModulus = 2,
    kVolumeConserving = 3,   // VolumeConservingLP (pf/volumeConserving.go)
    kSpectralViscosity = 4,  // SpectralViscosity (pf/spectralViscosity.go)
    kTensorialHessian = 5,   // TensorialHessian (pf/tensorialHessian.go)
    kConservativeNoise = 6,  // ConservativeNoise (pf/noise.go)
    kNumTermKinds = 7
};

struct TermDesc {
    int kind = kConcrete;
    std::complex<double> coef{1.0, 0.0};  // concrete: sign * prod Pow(scalar, p); user: +-1
    int brick = -1;                       // concrete: brick index (fields then derived), -1 none
    bool has_lap = false;
    int lap = 0;                          // concrete: LaplacianN power; user: total LAP prefix power
    int user = -1;                        // index into the user-term table of its kind
    double sign = 1.0;                    // concrete: the sign of the term
    std::vector<std::pair<std::string, double>> factors;  // concrete: (scalar, power) factors of coef
};

// d coef / d scalar of a concrete term (real scalars): sign * sum over the factors
// of the scalar of p s^(p-1) times the other factors
double coef_derivative(const TermDesc& t, const std::map<std::string, std::complex<double>>& scalars,
                       const std::string& name);

struct DerivedDesc {
    std::string name;
    int kind = 0;  // 0: product of cmplx.Pow factors (util.go:38-65); 1: polynomial of real parts;
                   // 2: Gaussian noise of standard deviation noise_std, redrawn at every sync
    double noise_std = 0.0;
    unsigned long long noise_seed = 0;
    // terms: coef * prod field^power
    std::vector<std::pair<double, std::vector<std::pair<int, double>>>> terms;
};

struct UserTerm {
    int kind = kSquaredGradient;
    std::string field;       // SquaredGradient.Field / HomogeneousModulusLinElast.FieldName
    double factor = 1.0;     // SquaredGradient.Factor
    double C[81] = {};       // elastic tensor C_ijkl, index i*27 + j*9 + k*3 + l
    double misfit[9] = {};   // misfit strain, row-major 3x3
    std::string indicator;   // VolumeConservingLP.Indicator (field = VolumeConservingLP.Field)
    double dt = 0.0;         // VolumeConservingLP.Dt
    long num_nodes = 0;      // VolumeConservingLP.NumNodes
    double eps = 0.0;        // SpectralViscosity.Eps
    double threshold = 0.0;  // SpectralViscosity.DissipationThreshold
    int power = 0;           // SpectralViscosity.Power
    double hess[9] = {};     // TensorialHessian.K, row-major d x d
    int hess_len = 0;        // len(K): 4 (d = 2) or 9 (d = 3)
    int cn_dim = 0;          // ConservativeNoise.Dim; the current fields are derived fields
    std::string cn_names[3];
};

struct SubStringDelimiter {
    std::string sub;
    std::string delim;
};

class Model {
public:
    // fields (pf.Field): name and complex128 data
    std::vector<std::string> field_names;
    std::vector<std::vector<std::complex<double>>> field_data;
    std::vector<DerivedDesc> derived;
    std::map<std::string, std::complex<double>> scalars;
    std::map<std::string, UserTerm> implicit_terms, explicit_terms;
    std::vector<std::string> equations;
    std::vector<std::pair<std::vector<TermDesc>, std::vector<TermDesc>>> rhs;
    std::vector<UserTerm> user_table[kNumTermKinds];

    void add_field(const std::string& name, std::vector<std::complex<double>> data);
    void add_scalar(const std::string& name, std::complex<double> v);
    void register_explicit_term(const std::string& name, const UserTerm& t);
    void register_implicit_term(const std::string& name, const UserTerm& t);
    void register_polynomial(const std::string& name,
                             const std::vector<std::pair<double, std::vector<double>>>& terms);
    // RegisterFunction(name, WhiteNoise{strength}.Generate) (noise.go:14-22)
    void register_white_noise(const std::string& name, double strength, unsigned long long seed);
    // NewConservativeNoise + RegisterExplicitTerm(term, &noise, noise.RequiredDerivedFields(N))
    // (noise.go:24-100): dim derived current fields "<prefix>_current_<comp>"
    void register_conservative_noise(const std::string& term, double strength, int dim, unsigned long long seed);
    void add_equation(const std::string& eq);
    void init();

    std::vector<std::string> all_field_names() const;
    bool is_field_name(const std::string& n) const;
    bool is_brick_name(const std::string& n) const;
    int brick_index(const std::string& n) const;
    long num_nodes() const;

    // parser pieces, exposed for tests through the C ABI
    std::pair<std::vector<TermDesc>, std::vector<TermDesc>> build(const std::string& eq);
    TermDesc concrete_term(const SubStringDelimiter& t);

private:
    void update_derived_fields(const std::string& eq);
    DerivedDesc derived_from_desc(const std::string& desc) const;
    bool valid_name(const std::string& name) const;
    TermDesc user_term(const UserTerm& u, const std::vector<std::string>& prefixes);
};

// pf/util.go / rhsBuilder.go helpers (exported for the parser tests)
std::vector<SubStringDelimiter> split_on_many(const std::string& value, const std::vector<std::string>& delims);
std::string field_name_from_leibniz(const std::string& leibniz);
bool is_bilinear(const std::string& term, const std::string& field, const std::vector<std::string>& names);
std::string remove_known_prefixes(std::string s);
std::vector<std::string> get_known_prefixes(std::string s);
std::string sort_factors(const std::string& expr);
std::string get_field_name(const std::string& term, const std::vector<std::string>& names);
double get_power(const std::string& pattern);
std::complex<double> go_cpow(std::complex<double> x, double p);
std::complex<double> go_cmul(std::complex<double> x, std::complex<double> y);

}  // namespace gpf
