// Asynchronous output staging: thermo rows and trajectory frames leave the
// compute stream through pinned host slots and are written by one writer
// thread, so an output step costs the compute stream a pack kernel and two
// enqueued copies instead of a host round trip. Counterpart of the
// reference's pinned staging rings and background writer
// (thermo_output.cu:151-352, trajectory_output.cu:313-436,
// host_writer.cpp:18-105, output_system.cu:91-284).
//
// Per output step the engine takes a free slot (acquire), enqueues on its
// compute stream whatever fills the slot's device area (the packed trajectory
// records), then submit() snapshots the control block into the slot, records
// an event on the compute stream, and lets a non-blocking copy stream move the
// slot to pinned host memory once that event has fired. The writer thread
// waits for the copy, hands the host copy to the sink, and frees the slot.
// Slots are reused only after the sink has returned, so a device area is never
// overwritten while its copy is pending. A sink error stops all further
// writing; acquire() and drain() rethrow it with its status.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "device.hpp"

namespace beads_amd {

class AsyncOutput {
 public:
  // sink(step, thermo, trajectory, control block, packed records or nullptr)
  using Sink = std::function<void(uint64_t, bool, bool, const Ctrl&, const unsigned char*)>;

  AsyncOutput(size_t record_bytes, Sink sink);
  ~AsyncOutput();
  AsyncOutput(const AsyncOutput&) = delete;
  AsyncOutput& operator=(const AsyncOutput&) = delete;

  // A free slot (waits for the writer when every slot is in flight).
  int acquire();
  // Device area for the packed records of a slot (record_bytes bytes).
  unsigned char* device_records(int slot);
  // Stage the slot: control block snapshot + copies after the work already
  // enqueued on `compute`; the writer handles it in submission order.
  void submit(int slot, const Ctrl* ctrl_dev, hipStream_t compute, uint64_t step, bool thermo,
              bool trajectory);
  // Give back a slot that was acquired but not submitted.
  void cancel(int slot);
  // Wait until everything submitted has been written.
  void drain();

 private:
  static constexpr int kSlots = 3;
  struct Slot {
    DeviceBuffer<unsigned char> device;  // [Ctrl | records]
    PinnedBuffer<unsigned char> host;
    hipEvent_t staged = nullptr;
    hipEvent_t copied = nullptr;
    bool busy = false;
  };
  struct Job {
    int slot;
    uint64_t step;
    bool thermo, trajectory;
  };
  static constexpr size_t kCtrlArea = 256;  // the records start 256-byte aligned
  static_assert(sizeof(Ctrl) <= kCtrlArea, "control block larger than its staging area");

  void writer_loop();
  void rethrow_locked();

  size_t record_bytes_;
  Sink sink_;
  hipStream_t copy_stream_ = nullptr;
  Slot slots_[kSlots];
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<Job> queue_;
  bool stop_ = false;
  bool failed_ = false;
  beads_status_t error_status_ = BEADS_OK;
  std::string error_message_;
  std::thread writer_;
};

}  // namespace beads_amd
