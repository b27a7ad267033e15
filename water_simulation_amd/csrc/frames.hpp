// frames.hpp -- asynchronous frame_%04d.bin output (see frames.cpp).
#pragma once

#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>

#include "common.hpp"

namespace wsim {

class FrameWriter {
public:
    FrameWriter(std::string dir, size_t max_points);
    ~FrameWriter();  // writes what was pushed, then stops
    FrameWriter(const FrameWriter&) = delete;
    FrameWriter& operator=(const FrameWriter&) = delete;

    // Frame `index` = the n points (x, y, z floats, device memory) as they are
    // once the work enqueued on `stream` so far has run.  Blocks only while
    // both buffers still hold frames not yet on disk.
    void push(const float* xyz, size_t n, int index, hipStream_t stream);
    // Waits until every pushed frame is written; reports a write error.
    void finish();

private:
    struct Slot {
        DeviceBuffer<float> dev;
        PinnedBuffer<float> host;
        hipEvent_t snapped = nullptr, copied = nullptr;
        size_t n = 0;
        int index = 0;
        bool busy = false;
    };
    void run();

    std::string dir_;
    size_t max_;
    hipStream_t copy_ = nullptr;
    Slot slots_[2];
    int next_ = 0;
    std::thread thread_;
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<int> queue_;
    bool stop_ = false;
    std::string error_;
};

}  // namespace wsim
