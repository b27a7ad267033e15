// frames.cpp -- asynchronous frame output for the scene driver.
//
// sim_cli writes frame_%04d.bin after each frame with a synchronous device
// to host copy and a file write on the simulation thread
// (apps/sim_cli/main.cpp:163-166, 191-193).  Here a frame costs the
// simulation stream one device-side snapshot of the positions; the copy to
// one of two pinned host buffers runs on a copy stream and the file is
// written by a background thread while the next frame is simulated.
#include "frames.hpp"

#include <cstdio>

#include "host_surface.hpp"

namespace wsim {

FrameWriter::FrameWriter(std::string dir, size_t max_points) : dir_(std::move(dir)), max_(max_points) {
    if (max_points == 0) fail(kInvalidArgument, "FrameWriter: max_points must be positive");
    WSIM_HIP(hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking));
    for (Slot& s : slots_) {
        s.dev = DeviceBuffer<float>(3 * max_points);
        s.host = PinnedBuffer<float>(3 * max_points);
        WSIM_HIP(hipEventCreateWithFlags(&s.snapped, hipEventDisableTiming));
        WSIM_HIP(hipEventCreateWithFlags(&s.copied, hipEventDisableTiming));
    }
    thread_ = std::thread([this] { run(); });
}

FrameWriter::~FrameWriter() {
    {
        std::lock_guard<std::mutex> g(m_);
        stop_ = true;
    }
    cv_.notify_all();
    if (thread_.joinable()) thread_.join();
    for (Slot& s : slots_) {
        (void)hipEventDestroy(s.snapped);
        (void)hipEventDestroy(s.copied);
    }
    (void)hipStreamDestroy(copy_);
}

void FrameWriter::push(const float* xyz, size_t n, int index, hipStream_t stream) {
    if (n > max_) fail(kInvalidArgument, "FrameWriter: more points than max_points");
    const int k = next_;
    next_ ^= 1;
    Slot& s = slots_[k];
    {
        // the slot's previous frame must be on disk before its buffers are reused
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [&] { return !s.busy; });
        if (!error_.empty()) fail(kRuntimeError, error_);
        s.busy = true;
    }
    s.n = n;
    s.index = index;
    if (n) {
        WSIM_HIP(hipMemcpyAsync(s.dev.data(), xyz, 12 * n, hipMemcpyDeviceToDevice, stream));
        WSIM_HIP(hipEventRecord(s.snapped, stream));
        WSIM_HIP(hipStreamWaitEvent(copy_, s.snapped, 0));
        WSIM_HIP(hipMemcpyAsync(s.host.data(), s.dev.data(), 12 * n, hipMemcpyDeviceToHost, copy_));
    }
    WSIM_HIP(hipEventRecord(s.copied, copy_));
    {
        std::lock_guard<std::mutex> g(m_);
        queue_.push_back(k);
    }
    cv_.notify_all();
}

void FrameWriter::finish() {
    std::unique_lock<std::mutex> g(m_);
    cv_.wait(g, [&] { return queue_.empty() && !slots_[0].busy && !slots_[1].busy; });
    if (!error_.empty()) fail(kRuntimeError, error_);
}

void FrameWriter::run() {
    for (;;) {
        int k;
        {
            std::unique_lock<std::mutex> g(m_);
            cv_.wait(g, [&] { return stop_ || !queue_.empty(); });
            if (queue_.empty()) return;  // stopping with nothing left
            k = queue_.front();
        }
        Slot& s = slots_[k];
        std::string err;
        if (hipEventSynchronize(s.copied) != hipSuccess) {
            err = "FrameWriter: device to host copy failed";
        } else {
            char name[32];
            std::snprintf(name, sizeof(name), "frame_%04d.bin", s.index);
            try {
                write_frame(dir_ + "/" + name, s.host.data(), s.n);
            } catch (const std::exception& e) {
                err = e.what();
            }
        }
        {
            std::lock_guard<std::mutex> g(m_);
            queue_.pop_front();
            s.busy = false;
            if (!err.empty() && error_.empty()) error_ = err;
        }
        cv_.notify_all();
    }
}

}  // namespace wsim
