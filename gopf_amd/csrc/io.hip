// io.hip -- field output in the reference's file formats.
//
//   k_pack_be64   Float64IO / Field.SaveReal / SaveFloat64 bytes (fileIO.go:57-62,
//                 model.go:35-41, fileIO.go:87-95): the real part as big-endian float64
//   k_minmax      MinReal / MaxReal (pfutil/sliceOperations.go:61-80)
//   k_pack_u8     Uint8IO / RealPartAsUint8 (fileIO.go:28-42, util.go:108-117)
//
// Host side: the XDMF text of CreateXDMF as Go's xml.Encoder writes it, SaveCsv, and the
// asynchronous snapshot writer.
#include "io.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>

#include "../../include/gpf.h"
#include "common.h"

namespace gpf {

namespace {

constexpr int kIoThreads = 256;

inline unsigned grid_for(long n) {
    long g = (n + kIoThreads - 1) / kIoThreads;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    return (unsigned)g;
}

template <typename T>
__global__ void __launch_bounds__(kIoThreads) k_pack_be64(const T* __restrict__ x, const double* __restrict__ mean,
                                                           uint64_t* __restrict__ out, long n) {
    const double mu = mean ? mean[0] : 0.0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += stride) {
        const double v = (double)x[i] + mu;
        out[i] = __builtin_bswap64((uint64_t)__double_as_longlong(v));
    }
}

// the reference's loops start from element 0 and replace on a strict comparison, so NaN
// elements are skipped unless element 0 is NaN (then the result is NaN); starting every
// thread from element 0 and combining partials the same way gives that result
template <typename T>
__global__ void __launch_bounds__(kIoThreads) k_minmax(const T* __restrict__ x, const double* __restrict__ mean,
                                                        long n, double* __restrict__ partial) {
    const double mu = mean ? mean[0] : 0.0;
    const double x0 = (double)x[0] + mu;
    double mn = x0, mx = x0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += stride) {
        const double v = (double)x[i] + mu;
        if (v < mn) mn = v;
        if (v > mx) mx = v;
    }
    __shared__ double smn[kIoThreads], smx[kIoThreads];
    smn[threadIdx.x] = mn;
    smx[threadIdx.x] = mx;
    __syncthreads();
    for (int s = kIoThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            if (smn[threadIdx.x + s] < smn[threadIdx.x]) smn[threadIdx.x] = smn[threadIdx.x + s];
            if (smx[threadIdx.x + s] > smx[threadIdx.x]) smx[threadIdx.x] = smx[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = smn[0];
        partial[2 * blockIdx.x + 1] = smx[0];
    }
}

__global__ void __launch_bounds__(kIoThreads) k_minmax_final(const double* __restrict__ partial, long nb,
                                                              double* __restrict__ mm) {
    if (threadIdx.x != 0) return;
    double mn = partial[0], mx = partial[1];
    for (long b = 1; b < nb; ++b) {
        if (partial[2 * b] < mn) mn = partial[2 * b];
        if (partial[2 * b + 1] > mx) mx = partial[2 * b + 1];
    }
    mm[0] = mn;
    mm[1] = mx;
}

// Go's float64 -> uint8 conversion truncates toward zero (a NaN gives 0 on amd64)
template <typename T>
__global__ void __launch_bounds__(kIoThreads) k_pack_u8(const T* __restrict__ x, const double* __restrict__ mean,
                                                         const double* __restrict__ mm, uint8_t* __restrict__ out,
                                                         long n) {
    const double mu = mean ? mean[0] : 0.0;
    const double mn = mm[0];
    double mx = mm[1];
    if (fabs(mx - mn) < 1e-10) mx = mn + 1.0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += stride) {
        const double v = (double)x[i] + mu;
        const double t = 255.0 * (v - mn) / (mx - mn);
        const long long q = (t == t && fabs(t) < 9.2e18) ? (long long)t : 0;
        out[i] = (uint8_t)(q & 0xff);
    }
}

}  // namespace

long io_blocks(long n) { return (long)grid_for(n); }

template <typename T>
hipError_t launch_pack_be64(const T* x, const double* mean, uint64_t* out, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_pack_be64<T>, dim3(grid_for(n)), dim3(kIoThreads), 0, s, x, mean, out, n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_minmax(const T* x, const double* mean, long n, double* partial, double* mm, hipStream_t s) {
    if (n <= 0) return hipErrorInvalidValue;
    const unsigned nb = grid_for(n);
    hipLaunchKernelGGL(k_minmax<T>, dim3(nb), dim3(kIoThreads), 0, s, x, mean, n, partial);
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(64), 0, s, partial, (long)nb, mm);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pack_u8(const T* x, const double* mean, const double* mm, uint8_t* out, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_pack_u8<T>, dim3(grid_for(n)), dim3(kIoThreads), 0, s, x, mean, mm, out, n);
    return hipGetLastError();
}

#define GPF_INST(T)                                                                                     \
    template hipError_t launch_pack_be64<T>(const T*, const double*, uint64_t*, long, hipStream_t);    \
    template hipError_t launch_minmax<T>(const T*, const double*, long, double*, double*, hipStream_t); \
    template hipError_t launch_pack_u8<T>(const T*, const double*, const double*, uint8_t*, long, hipStream_t);
GPF_INST(float)
GPF_INST(double)
#undef GPF_INST

// ---------------------------------------------------------------------------
// XDMF (pf/xdmf.go)
// ---------------------------------------------------------------------------
namespace {

// Go's xml escaping of attribute values and character data (EscapeText)
std::string xml_escape(const std::string& v) {
    std::string o;
    for (char c : v) {
        switch (c) {
            case '"': o += "&#34;"; break;
            case '\'': o += "&#39;"; break;
            case '&': o += "&amp;"; break;
            case '<': o += "&lt;"; break;
            case '>': o += "&gt;"; break;
            case '\t': o += "&#x9;"; break;
            case '\n': o += "&#xA;"; break;
            case '\r': o += "&#xD;"; break;
            default: o += c;
        }
    }
    return o;
}

// the indentation rule of Go's xml printer (writeIndent): a start tag begins a new line
// at the current depth; an end tag does too unless nothing but character data was written
// since its start tag
struct XmlOut {
    std::string s, indent;
    int depth = 0;
    bool indented_in = false, put_newline = false;
    void nl(int delta) {
        if (indent.empty()) return;
        if (delta < 0) {
            --depth;
            if (indented_in) {
                indented_in = false;
                return;
            }
            indented_in = false;
        }
        if (put_newline) s += '\n';
        else put_newline = true;
        for (int i = 0; i < depth; ++i) s += indent;
        if (delta > 0) {
            ++depth;
            indented_in = true;
        }
    }
    // attributes with an empty value are omitted (every attribute here is omitempty
    // except TimeType, which is never empty)
    void start(const char* name, std::initializer_list<std::pair<const char*, std::string>> attrs) {
        nl(1);
        s += '<';
        s += name;
        for (const auto& a : attrs) {
            if (a.second.empty()) continue;
            s += ' ';
            s += a.first;
            s += "=\"";
            s += xml_escape(a.second);
            s += '"';
        }
        s += '>';
    }
    void text(const std::string& t) { s += xml_escape(t); }
    void end(const char* name) {
        nl(-1);
        s += "</";
        s += name;
        s += '>';
    }
};

std::string itoa_nz(int v) { return v ? std::to_string(v) : std::string(); }

// XDMFDataItem attribute order: Dimension, Format, NumberType, Dimensions, Precision, DataType, Endian
void data_item(XmlOut& o, int dim, const std::string& format, const std::string& value, const std::string& number_type,
               const std::string& dimensions, int precision, const std::string& data_type, const std::string& endian) {
    o.start("DataItem", {{"Dimension", itoa_nz(dim)},
                         {"Format", format},
                         {"NumberType", number_type},
                         {"Dimensions", dimensions},
                         {"Precision", itoa_nz(precision)},
                         {"DataType", data_type},
                         {"Endian", endian}});
    o.text(value);
    o.end("DataItem");
}

std::string go_f(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "+Inf" : "-Inf";
    char b[512];
    std::snprintf(b, sizeof(b), "%f", v);
    return b;
}

}  // namespace

std::string xdmf_text_indent(const std::vector<std::string>& fields, const std::string& prefix, int num,
                             const std::vector<int>& dims, const std::string& indent) {
    const int dim = (int)dims.size();
    std::string dimensions, geo_type;
    if (dim == 2) {
        dimensions = std::to_string(dims[0]) + " " + std::to_string(dims[1]);
        geo_type = "ORIGIN_DXDY";
    } else if (dim == 3) {
        dimensions = std::to_string(dims[0]) + " " + std::to_string(dims[1]) + " " + std::to_string(dims[2]);
        geo_type = "ORIGIN_DXDYDZ";
    } else {
        throw Error(GPF_EINVAL, "Length of domain size to be either 2 or 3");
    }
    if (num < 0) throw Error(GPF_EINVAL, "negative number of time steps");
    XmlOut o;
    o.indent = indent;
    o.start("Xdmf", {});
    o.start("Domain", {});
    o.start("Topology", {{"name", "topo"}, {"TopologyType", std::to_string(dim) + "DCoRectMesh"},
                         {"Dimensions", dimensions}, {"Reference", ""}});
    o.end("Topology");
    o.start("Geometry", {{"name", "geo"}, {"Type", geo_type}, {"Reference", ""}});
    data_item(o, 0, "XML", "0.0 0.0 0.0", "", std::to_string(dim), 0, "", "");
    data_item(o, 0, "XML", "1.0 1.0 1.0", "", std::to_string(dim), 0, "", "");
    o.end("Geometry");
    o.start("Grid", {{"Name", "TimeSeries"}, {"GridType", "Collection"}, {"CollectionType", "Temporal"}});
    for (int i = 0; i < num; ++i) {
        o.start("Grid", {{"Name", "T" + std::to_string(i + 1)}, {"GridType", "Uniform"}, {"CollectionType", ""}});
        for (const auto& f : fields) {
            o.start("Attribute", {{"Name", f}, {"Center", "Node"}});
            data_item(o, 0, "Binary", prefix + "_" + f + "_" + std::to_string(i) + ".bin", "", dimensions, 8, "Float",
                      "Big");
            o.end("Attribute");
        }
        o.start("Topology", {{"Reference", "/Xdmf/Domain/Topology[1]"}});
        o.end("Topology");
        o.start("Geometry", {{"Reference", "/Xdmf/Domain/Geometry[1]"}});
        o.end("Geometry");
        o.end("Grid");
    }
    // the collection's own (zero) Topology and Geometry: Go writes struct-valued elements
    // even when tagged omitempty
    o.start("Topology", {});
    o.end("Topology");
    o.start("Geometry", {});
    o.end("Geometry");
    std::string times;
    for (int i = 0; i < num; ++i) times += go_f((double)i / (double)num) + " ";
    times += std::to_string(num);
    o.start("Time", {{"TimeType", "HyperSlab"}});
    data_item(o, 0, "XML", times, "Float", std::to_string(dim), 0, "", "");
    o.end("Time");
    o.end("Grid");
    o.end("Domain");
    o.end("Xdmf");
    return o.s;
}

std::string xdmf_text(const std::vector<std::string>& fields, const std::string& prefix, int num,
                      const std::vector<int>& dims) {
    return xdmf_text_indent(fields, prefix, num, dims, "    ");
}

// ---------------------------------------------------------------------------
// CSV and raw float64 files (pf/fileIO.go)
// ---------------------------------------------------------------------------
namespace {

// encoding/csv's fieldNeedsQuotes with Comma ','
bool csv_needs_quotes(const std::string& f) {
    if (f.empty()) return false;
    if (f == "\\.") return true;
    if (f.find_first_of(",\"\r\n") != std::string::npos) return true;
    const unsigned char c = (unsigned char)f[0];
    return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

void csv_field(std::string& line, const std::string& f) {
    if (!csv_needs_quotes(f)) {
        line += f;
        return;
    }
    line += '"';
    for (char c : f) {
        if (c == '"') line += "\"\"";
        else line += c;
    }
    line += '"';
}

}  // namespace

void write_csv(const std::string& path, const std::vector<std::string>& names,
               const std::vector<std::vector<double>>& values, const std::vector<int>& dims) {
    if (dims.empty()) throw Error(GPF_EINVAL, "Domain size not given. Data will not be written to file");
    if (values.empty()) {
        std::fprintf(stderr, "No data provided, will not write anything\n.");
        return;
    }
    if (dims.size() != 2 && dims.size() != 3) throw Error(GPF_EINVAL, "util: Domain size has to be either 2 or 3");
    const long n = (long)values[0].size();
    long prod = 1;
    for (int d : dims) prod *= d;
    if (prod != n)
        throw Error(GPF_EINVAL, "Inconsistent domain size. Expected " + std::to_string(n) + " nodes, got " +
                                    std::to_string(prod));
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw Error(GPF_EINVAL, "Could not open file: " + path);
    std::string line = "X,Y,Z";
    for (const auto& nm : names) {
        line += ',';
        csv_field(line, nm);
    }
    line += '\n';
    std::fwrite(line.data(), 1, line.size(), f);
    for (long i = 0; i < n; ++i) {
        long pos[3] = {0, 0, 0};
        if (dims.size() == 2) {
            pos[1] = i % dims[1];
            pos[0] = i / dims[1];
        } else {
            pos[1] = i % dims[1];
            pos[0] = (i / dims[1]) % dims[0];
            pos[2] = i / ((long)dims[0] * dims[1]);
        }
        line.clear();
        line += std::to_string(pos[0]) + "," + std::to_string(pos[1]) + "," + std::to_string(pos[2]);
        for (const auto& v : values) {
            line += ',';
            csv_field(line, go_f(v[i]));
        }
        line += '\n';
        std::fwrite(line.data(), 1, line.size(), f);
    }
    std::fclose(f);
}

void save_float64(const std::string& path, const double* data, long n) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw Error(GPF_EINVAL, "open " + path + ": cannot create file");
    std::vector<uint64_t> buf(4096);
    for (long i0 = 0; i0 < n; i0 += (long)buf.size()) {
        const long m = std::min((long)buf.size(), n - i0);
        for (long j = 0; j < m; ++j) {
            uint64_t u;
            std::memcpy(&u, data + i0 + j, 8);
            buf[j] = __builtin_bswap64(u);
        }
        std::fwrite(buf.data(), 8, m, f);
    }
    std::fclose(f);
}

std::vector<double> load_float64(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw Error(GPF_EINVAL, "open " + path + ": no such file");
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<double> out(size / 8);
    std::vector<uint64_t> raw(out.size());
    const size_t got = std::fread(raw.data(), 8, raw.size(), f);
    std::fclose(f);
    for (size_t i = 0; i < got; ++i) {
        const uint64_t u = __builtin_bswap64(raw[i]);
        std::memcpy(&out[i], &u, 8);
    }
    return out;
}

// ---------------------------------------------------------------------------
// asynchronous snapshot writer
// ---------------------------------------------------------------------------
SnapshotWriter::SnapshotWriter(int device) : device_(device) {
    HIPCHK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&packed_, hipEventDisableTiming));
    for (auto& s : slots_) HIPCHK(hipEventCreateWithFlags(&s.copied, hipEventDisableTiming));
    th_ = std::thread([this] { run(); });
}

SnapshotWriter::~SnapshotWriter() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (th_.joinable()) th_.join();
    (void)hipStreamSynchronize(st_);
    for (auto& s : slots_) {
        if (s.dev) (void)hipFree(s.dev);
        if (s.host) (void)hipHostFree(s.host);
        if (s.copied) (void)hipEventDestroy(s.copied);
    }
    if (packed_) (void)hipEventDestroy(packed_);
    if (st_) (void)hipStreamDestroy(st_);
}

int SnapshotWriter::acquire(size_t bytes) {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return !slots_[0].busy || !slots_[1].busy; });
    const int k = !slots_[0].busy ? 0 : 1;
    Slot& s = slots_[k];
    if (s.cap < bytes) {
        if (s.dev) (void)hipFree(s.dev);
        if (s.host) (void)hipHostFree(s.host);
        s.dev = s.host = nullptr;
        s.cap = 0;
        HIPCHK(hipMalloc(&s.dev, bytes));
        HIPCHK(hipHostMalloc(&s.host, bytes, hipHostMallocDefault));
        s.cap = bytes;
    }
    s.busy = true;
    return k;
}

void SnapshotWriter::submit(int slot, hipStream_t producer, size_t bytes,
                            std::vector<std::pair<std::string, std::pair<size_t, size_t>>> files) {
    try {
        HIPCHK(hipEventRecord(packed_, producer));
        HIPCHK(hipStreamWaitEvent(st_, packed_, 0));
        HIPCHK(hipMemcpyAsync(slots_[slot].host, slots_[slot].dev, bytes, hipMemcpyDeviceToHost, st_));
        HIPCHK(hipEventRecord(slots_[slot].copied, st_));
    } catch (...) {
        std::lock_guard<std::mutex> lk(mu_);
        slots_[slot].busy = false;
        cv_.notify_all();
        throw;
    }
    {
        std::lock_guard<std::mutex> lk(mu_);
        q_.push_back(Job{slot, std::move(files)});
    }
    cv_.notify_all();
}

void SnapshotWriter::flush() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return q_.empty() && !slots_[0].busy && !slots_[1].busy; });
    if (!error_.empty()) {
        const std::string e = error_;
        error_.clear();
        throw Error(GPF_EINVAL, e);
    }
}

void SnapshotWriter::run() {
    (void)hipSetDevice(device_);
    for (;;) {
        Job j;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || !q_.empty(); });
            if (q_.empty()) return;
            j = std::move(q_.front());
            q_.pop_front();
        }
        std::string err;
        const hipError_t e = hipEventSynchronize(slots_[j.slot].copied);
        if (e != hipSuccess) {
            err = std::string("snapshot copy: ") + hipGetErrorString(e);
        } else {
            const char* base = static_cast<const char*>(slots_[j.slot].host);
            for (const auto& f : j.files) {
                FILE* fp = std::fopen(f.first.c_str(), "wb");
                if (!fp) {
                    // Float64IO prints the error and goes on (fileIO.go:87-93)
                    std::fprintf(stderr, "open %s: cannot create file\n", f.first.c_str());
                    if (err.empty()) err = "open " + f.first + ": cannot create file";
                    continue;
                }
                const size_t w = std::fwrite(base + f.second.first, 1, f.second.second, fp);
                std::fclose(fp);
                if (w != f.second.second && err.empty()) err = "short write to " + f.first;
            }
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            slots_[j.slot].busy = false;
            if (!err.empty() && error_.empty()) error_ = err;
        }
        cv_.notify_all();
    }
}

}  // namespace gpf
