// Allocation-free neighbour-table build, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  It runs exactly the kernels of
// lsmps_build_neighbor_lists (cell lists of both particle sets, count pass,
// scan, capacity check, write pass), so its tables are bit-identical; only the
// scratch differs: the per-cell counts and cursors, per-particle cell labels,
// sorted point records and scan partials live in one static device array
// instead of the context's growable buffers.
//
// Limits: at most kNoAllocParticles fluid particles and as many wall particles
// per call, and at most kNoAllocCells cells.  Every GPU the library runs on
// holds its own copy of the array (about 44 MB) until the process ends.  All
// calls on one GPU share that copy: calls that overlap on one GPU (two
// streams, two threads) race and must be serialised by the caller.  Nothing in
// the array needs to start at zero: every counter is cleared by a fill kernel
// before use and every other entry is written before it is read.
#include <array>
#include <mutex>
#include <string>

#include "common.hpp"
#include "stages.hpp"

namespace lsmps {
namespace {

constexpr size_t kNoAllocParticles = size_t(1) << 20;  // per particle set
constexpr size_t kNoAllocCells = size_t(1) << 20;
constexpr size_t kNoAllocPartials = 4096;  // scan partials (n = 2^20 needs 531)

constexpr size_t align256(size_t bytes) { return (bytes + 255) / 256 * 256; }
constexpr size_t kCellBytes = align256((kNoAllocCells + 1) * sizeof(int32_t));
constexpr size_t kLabelBytes = align256(kNoAllocParticles * sizeof(int32_t));
constexpr size_t kPointBytes = align256(kNoAllocParticles * kSortedPointBytes);
constexpr size_t kPartialBytes = align256(kNoAllocPartials * sizeof(int32_t));
constexpr size_t kCursorOffset = kCellBytes;
constexpr size_t kLabelOffset = 2 * kCellBytes;
constexpr size_t kPointOffset = kLabelOffset + kLabelBytes;
constexpr size_t kPartialOffset = kPointOffset + kPointBytes;
constexpr size_t kScratchBytes = kPartialOffset + kPartialBytes;

// The scratch of the allocation-free build.  In host code this name is a
// host-side stand-in, not a device address: host code only ever uses the
// per-device address hipGetSymbolAddress returns (scratch_ptr below).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
std::array<unsigned char*, kMaxDevices> g_scratch_ptr{};

unsigned char* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

}  // namespace

void noalloc_prepare_device(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  require(device >= 0 && device < kMaxDevices,
          "device index beyond the allocation-free scratch table", LSMPS_ERR_DEVICE);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  LSMPS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // this device is current
  require(p != nullptr, "hipGetSymbolAddress returned no address for the static scratch",
          LSMPS_ERR_DEVICE);
  g_scratch_ptr[device] = static_cast<unsigned char*>(p);
}

void build_neighbor_lists_noalloc(Context& ctx, const lsmps_fluid_soa& fluid,
                                  const lsmps_wall_soa& walls, const lsmps_config& config,
                                  const lsmps_cell_list& fluid_cells,
                                  const lsmps_cell_list& wall_cells,
                                  const lsmps_neighbor_list& fluid_neighbors,
                                  const lsmps_neighbor_list& wall_neighbors) {
  unsigned char* base = scratch_ptr(ctx.device);
  require(base != nullptr, "allocation-free scratch not set up for this device",
          LSMPS_ERR_INTERNAL);
  require(config.cell_dims[0] > 0 && config.cell_dims[1] > 0 && config.cell_dims[2] > 0,
          "Invalid cell grid configuration");
  const size_t cells = static_cast<size_t>(config.cell_dims[0]) * config.cell_dims[1] *
                       config.cell_dims[2];
  auto limit = [](const char* what, size_t got, size_t max) {
    if (got > max)
      raise(LSMPS_ERR_CAPACITY, std::string("lsmps_build_neighbor_lists_noalloc: ") +
                                    std::to_string(got) + " " + what +
                                    " exceed the static scratch limit of " +
                                    std::to_string(max));
  };
  limit("fluid particles", fluid.count, kNoAllocParticles);
  limit("wall particles", walls.count, kNoAllocParticles);
  limit("cells", cells, kNoAllocCells);
  GeometryScratch sc;
  sc.counts = reinterpret_cast<int32_t*>(base);
  sc.cursor = reinterpret_cast<int32_t*>(base + kCursorOffset);
  sc.cells_capacity = kNoAllocCells + 1;
  sc.labels = reinterpret_cast<int32_t*>(base + kLabelOffset);
  sc.points = reinterpret_cast<char*>(base + kPointOffset);
  sc.particles_capacity = kNoAllocParticles;
  sc.partials = reinterpret_cast<int32_t*>(base + kPartialOffset);
  sc.partials_capacity = kNoAllocPartials;
  const AllocationFreeze frozen;  // any device allocation below raises
  build_neighbor_lists_with(ctx, sc, fluid, walls, config, fluid_cells, wall_cells,
                            fluid_neighbors, wall_neighbors, nullptr, nullptr);
}

}  // namespace lsmps
