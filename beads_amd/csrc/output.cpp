// AsyncOutput: pinned staging slots, a copy stream and one writer thread.
#include "output.hpp"

#include <exception>
#include <utility>

namespace beads_amd {

AsyncOutput::AsyncOutput(size_t record_bytes, Sink sink)
    : record_bytes_(record_bytes), sink_(std::move(sink)) {
  BEADS_HIP(hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking));
  for (Slot& s : slots_) {
    s.device.resize(kCtrlArea + record_bytes_);
    s.host.resize(kCtrlArea + record_bytes_);
    BEADS_HIP(hipEventCreateWithFlags(&s.staged, hipEventDisableTiming));
    BEADS_HIP(hipEventCreateWithFlags(&s.copied, hipEventDisableTiming));
  }
  writer_ = std::thread([this] { writer_loop(); });
}

AsyncOutput::~AsyncOutput() {
  {
    std::unique_lock<std::mutex> lock(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (writer_.joinable()) writer_.join();  // writes what was submitted, then ends
  if (copy_stream_ != nullptr) (void)hipStreamSynchronize(copy_stream_);
  for (Slot& s : slots_) {
    if (s.staged != nullptr) (void)hipEventDestroy(s.staged);
    if (s.copied != nullptr) (void)hipEventDestroy(s.copied);
  }
  if (copy_stream_ != nullptr) (void)hipStreamDestroy(copy_stream_);
}

void AsyncOutput::rethrow_locked() {
  if (failed_) fail(error_status_, error_message_);
}

int AsyncOutput::acquire() {
  std::unique_lock<std::mutex> lock(mu_);
  for (;;) {
    rethrow_locked();
    for (int k = 0; k < kSlots; ++k) {
      if (!slots_[k].busy) {
        slots_[k].busy = true;
        return k;
      }
    }
    cv_.wait(lock);
  }
}

unsigned char* AsyncOutput::device_records(int slot) {
  return slots_[slot].device.data() + kCtrlArea;
}

void AsyncOutput::submit(int slot, const Ctrl* ctrl_dev, hipStream_t compute, uint64_t step,
                         bool thermo, bool trajectory) {
  Slot& s = slots_[slot];
  BEADS_HIP(hipMemcpyAsync(s.device.data(), ctrl_dev, sizeof(Ctrl), hipMemcpyDeviceToDevice,
                           compute));
  BEADS_HIP(hipEventRecord(s.staged, compute));
  BEADS_HIP(hipStreamWaitEvent(copy_stream_, s.staged, 0));
  const size_t bytes = kCtrlArea + (trajectory ? record_bytes_ : 0);
  BEADS_HIP(hipMemcpyAsync(s.host.data(), s.device.data(), bytes, hipMemcpyDeviceToHost,
                           copy_stream_));
  BEADS_HIP(hipEventRecord(s.copied, copy_stream_));
  {
    std::unique_lock<std::mutex> lock(mu_);
    queue_.push_back(Job{slot, step, thermo, trajectory});
  }
  cv_.notify_all();
}

void AsyncOutput::cancel(int slot) {
  {
    std::unique_lock<std::mutex> lock(mu_);
    slots_[slot].busy = false;
  }
  cv_.notify_all();
}

void AsyncOutput::drain() {
  std::unique_lock<std::mutex> lock(mu_);
  for (;;) {
    rethrow_locked();
    bool idle = queue_.empty();
    for (const Slot& s : slots_) idle = idle && !s.busy;
    if (idle) return;
    cv_.wait(lock);
  }
}

void AsyncOutput::writer_loop() {
  for (;;) {
    Job job{};
    {
      std::unique_lock<std::mutex> lock(mu_);
      cv_.wait(lock, [this] { return stop_ || !queue_.empty(); });
      if (queue_.empty()) return;  // stop_ and nothing left
      job = queue_.front();
      queue_.pop_front();
    }
    Slot& s = slots_[job.slot];
    bool ok = true;
    beads_status_t status = BEADS_OK;
    std::string message;
    try {
      BEADS_HIP(hipEventSynchronize(s.copied));
      bool skip;
      {
        std::unique_lock<std::mutex> lock(mu_);
        skip = failed_;  // after an error nothing more is written
      }
      if (!skip) {
        const Ctrl* c = reinterpret_cast<const Ctrl*>(s.host.data());
        sink_(job.step, job.thermo, job.trajectory, *c,
              job.trajectory ? s.host.data() + kCtrlArea : nullptr);
      }
    } catch (const Error& e) {
      ok = false;
      status = e.status();
      message = e.what();
    } catch (const std::exception& e) {
      ok = false;
      status = BEADS_E_RUNTIME;
      message = e.what();
    } catch (...) {
      ok = false;
      status = BEADS_E_RUNTIME;
      message = "unknown output writer error";
    }
    {
      std::unique_lock<std::mutex> lock(mu_);
      if (!ok && !failed_) {
        failed_ = true;
        error_status_ = status;
        error_message_ = message;
      }
      s.busy = false;
    }
    cv_.notify_all();
  }
}

}  // namespace beads_amd
