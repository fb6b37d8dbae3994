// io.h -- field output in the reference's file formats (pf/fileIO.go, pf/xdmf.go).
//
// Snapshots of resident fields are packed into their file bytes on the device
// (big-endian float64, or uint8 quantised between the field's min and max), moved to
// pinned host memory by an asynchronous copy on a stream of their own, and written by a
// host thread, so the solver's stream keeps stepping while a snapshot drains.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace gpf {

constexpr int kIoFloat64 = 0;  // Float64IO
constexpr int kIoUint8 = 1;    // Uint8IO

// out[i] = big-endian bytes of (double)x[i] + *mean  (mean may be null)
template <typename T>
hipError_t launch_pack_be64(const T* x, const double* mean, uint64_t* out, long n, hipStream_t s);
// mm[0] = min_i, mm[1] = max_i of (double)x[i] + *mean (partial: 2 * io_blocks(n) doubles)
template <typename T>
hipError_t launch_minmax(const T* x, const double* mean, long n, double* partial, double* mm, hipStream_t s);
// out[i] = uint8(255 (v - min) / (max - min)), v = (double)x[i] + *mean, with max = min + 1
// when |max - min| < 1e-10 (RealPartAsUint8, pf/util.go:108-117)
template <typename T>
hipError_t launch_pack_u8(const T* x, const double* mean, const double* mm, uint8_t* out, long n, hipStream_t s);
long io_blocks(long n);

// CreateXDMF marshalled as the reference's xml.Encoder with a four-space indent does
// (pf/xdmf.go:83-173); dims has 2 or 3 entries
std::string xdmf_text(const std::vector<std::string>& fields, const std::string& prefix, int num,
                      const std::vector<int>& dims);
// the same with another indent ("" is xml.Encoder without Indent: no newlines)
std::string xdmf_text_indent(const std::vector<std::string>& fields, const std::string& prefix, int num,
                             const std::vector<int>& dims, const std::string& indent);
// SaveCsv (pf/fileIO.go:147-200): header X,Y,Z,names; one record per node with the node's
// pfutil.Pos and each value as Go's %f
void write_csv(const std::string& path, const std::vector<std::string>& names,
               const std::vector<std::vector<double>>& values, const std::vector<int>& dims);
// SaveFloat64 / LoadFloat64 (pf/fileIO.go:65-95)
void save_float64(const std::string& path, const double* data, long n);
std::vector<double> load_float64(const std::string& path);

// asynchronous writer of device-packed snapshots (double-buffered)
class SnapshotWriter {
public:
    explicit SnapshotWriter(int device);
    ~SnapshotWriter();
    SnapshotWriter(const SnapshotWriter&) = delete;
    SnapshotWriter& operator=(const SnapshotWriter&) = delete;

    // a free slot with at least `bytes` of device and pinned host memory; waits for the
    // oldest snapshot in flight when both slots are busy
    int acquire(size_t bytes);
    void* device_buffer(int slot) const { return slots_[slot].dev; }
    // after the packing kernels were enqueued on `producer`: copy the slot to the host on
    // the writer's stream and queue the files (path, byte offset, byte count)
    void submit(int slot, hipStream_t producer, size_t bytes,
                std::vector<std::pair<std::string, std::pair<size_t, size_t>>> files);
    // waits until every queued file is written; throws the first write error
    void flush();

private:
    struct Slot {
        void* dev = nullptr;
        void* host = nullptr;
        size_t cap = 0;
        hipEvent_t copied = nullptr;
        bool busy = false;
    };
    struct Job {
        int slot;
        std::vector<std::pair<std::string, std::pair<size_t, size_t>>> files;
    };
    void run();
    int device_;
    hipStream_t st_ = nullptr;
    hipEvent_t packed_ = nullptr;
    Slot slots_[2];
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Job> q_;
    bool stop_ = false;
    std::string error_;
    std::thread th_;
};

}  // namespace gpf
