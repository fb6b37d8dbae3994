// batch.h -- batched single-field simulations (see batch.hip).
#pragma once

#include <string>
#include <vector>

#include "../../include/gpf.h"
#include "common.h"
#include "model.h"
#include "spectral.h"

namespace gpf {

using BatchError = Error;

// structure of a single-field model the fused path can run
struct FusedShape {
    std::vector<double> gpoly;  // explicit source g(c) = sum_q gpoly[q] c^q
    int src_lap = 0;            // LAP power of the source
    bool has_src = false;
    KPoly denom, lin;           // implicit D(k) and linear explicit E(k)
};

// classify a parsed model for the fused path (engine.hip); false with a reason otherwise
bool fused_shape(const Model& m, FusedShape* out, std::string* why);
// same powers / term counts (coefficients may differ)
bool same_structure(const FusedShape& a, const FusedShape& b);

class BatchBase {
public:
    virtual ~BatchBase() = default;
    virtual int nsims() const = 0;
    virtual void set_scalar(int sim, const std::string& name, double v) = 0;
    virtual void set_field(int sim, const double* in, long n) = 0;
    virtual void get_field(int sim, double* out, long n) = 0;
    virtual void propagate(int nsteps) = 0;
    virtual void synchronize() = 0;
    virtual int tuning(int* out, int maxn) = 0;
    virtual void set_layout(const int* v, int n) = 0;
};

BatchBase* make_batch(const Model& m, int ncls, const int* shapes, const int* counts, double dt, const gpf_options& o);

}  // namespace gpf
