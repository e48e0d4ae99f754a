// itsca_runtime.hip -- context, HBM buffers and the C ABI of libitsca_hip.so.
//
// Replaces the reference's CUDA runtime layer (gpu/Vector.h, gpu/Context.h,
// gpu/Stream.h, src/CA/gpu/cuda/{Context,Stream,Utils}.cu) and its per-layer
// host round trips (Tracker.cu:285-473): one upload per batch, a device-
// resident pipeline with no host synchronisation inside a run, and one small
// read-back of the output sizes per run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/itsca.h"
#include "itsca_device.h"
#include "itsca_kernels.h"
#include "itsca_mlp.h"
#include "itsca_mlp_kernels.h"

using namespace itsca;

namespace {

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf()
  {
    if (p) (void)hipFree(p);
  }
  // Grow-only; contents are not preserved.
  hipError_t reserve(size_t want)
  {
    if (want <= n && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    const size_t cnt = std::max<size_t>(want, 1);
    hipError_t e = hipMalloc(&p, cnt * sizeof(T));
    if (e == hipSuccess) n = cnt;
    return e;
  }
};

// Pinned (page-locked) host staging, grow-only, contents kept on growth:
// host-device copies from and to it run as direct DMA at the link's rate,
// where pageable memory goes through the runtime's bounce buffers.  A
// growth frees the old block, so no copy may be in flight from it then (the
// entry points that fill it synchronize their stream before returning).
template <typename T>
struct PinnedBuf {
  T* p = nullptr;
  size_t n = 0, cap = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf()
  {
    if (p) (void)hipHostFree(p);
  }
  hipError_t reserve(size_t want)
  {
    if (want <= cap && p) return hipSuccess;
    const size_t c = std::max<size_t>({want, cap + cap / 2, (size_t)4096});
    T* q = nullptr;
    const hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&q), c * sizeof(T), hipHostMallocDefault);
    if (e != hipSuccess) return e;
    if (n) std::memcpy(q, p, n * sizeof(T));
    if (p) (void)hipHostFree(p);
    p = q;
    cap = c;
    return hipSuccess;
  }
  void clear() { n = 0; }
  size_t size() const { return n; }
  T* data() { return p; }
  const T* data() const { return p; }
};

struct HipError {
  hipError_t code;
  std::string where;
};

#define HIP_TRY(expr)                                   \
  do {                                                  \
    hipError_t e_ = (expr);                             \
    if (e_ != hipSuccess) throw HipError{e_, #expr};    \
  } while (0)

}  // namespace

// Largest output capacity of a stage (element indices are int32).
constexpr double kMaxCapacity = 2.0e9;

// ---------------------------------------------------------------------------
// Device scratch of itsca_run_noalloc: every intermediate array of one run
// (sorted clusters, tables, tracklets, cells, edges, roads, chunk plans) is
// carved out of this array, so the entry point allocates nothing.  Each GPU
// the library runs on holds its own copy until the process ends.  In host
// code `scratch` names a host-side stand-in, never device memory: its device
// address is taken once per device with hipGetSymbolAddress (scratchAddress)
// and only that address is used.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[ITSCA_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
void* gScratchAddr[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be current.
void* scratchAddress(int device)
{
  std::lock_guard<std::mutex> lock(gScratchMutex);
  if (device < 0 || device >= kMaxDevices) throw HipError{hipErrorInvalidDevice, "scratchAddress: device ordinal"};
  if (!gScratchAddr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) throw HipError{e, "hipGetSymbolAddress(scratch)"};
    gScratchAddr[device] = p;
  }
  return gScratchAddr[device];
}

}  // namespace

struct itsca_ctx {
  int device = 0;
  int cus = 0;
  int persistentBlocks = 0;
  int blockSize = 256;  // workgroup (and chunk) size of the stage kernels
  int zSub = kDefaultZSub;  // z sub-bins per bin of the tracklet search index
  int phiSub = 1;           // phi sub-rows per row of the tracklet search index
  double capacityScale = 1.0;
  hipStream_t stream = nullptr;
  std::string err;
  bool profiling = false;
  KernelTimer timer{};
  bool timerReady = false;
  bool ran = false;
  // learned tracklet filter
  bool filterOn = false;
  float threshold = 0.0f;
  float phiGate[kTrackletLayers] = {kPhiCut, kPhiCut, kPhiCut, kPhiCut, kPhiCut, kPhiCut};
  bool filterFused = false;  // itsca_set_filter(ctx, 2, ...)
  int mlpStep = 0;
  int trainN = 0, trainCap = 0;
  DBuf<float> mlpMaster, mlpMom1, mlpMom2, mlpWq, mlpPartial, mlpGrad, mlpLoss, trainScores;
  DBuf<uint16_t> mlpWbf, trainX;
  DBuf<uint8_t> trainY, trkLabel;
  // whole-pipeline HIP graph, captured once per batch shape and buffer set
  bool useGraph = false;
  bool graphValid = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;

  // staged host input (the cluster arrays in pinned memory)
  PinnedBuf<float> hx, hy, hz;
  PinnedBuf<int> hid, hmc;
  std::vector<float> hvert;
  std::vector<int> hLayerOff, hEvVert;  // hLayerOff: 8 per event; hEvVert: first vertex of each event

  // batch shape
  int W = 0, E = 0;
  std::vector<int> itemEvent;
  std::vector<int> recItem;  // the item whose cluster records each item reads
  std::vector<Seg> clSeg;  // 7 per item
  long sortedClusters = 0;
  int tCap = 0, cCap = 0, eCap = 0, rCap = 0;
  int maxSegs = 0, maxChunks = 0;

  // results of the last run
  std::vector<Seg> trkSeg, celSeg, edgSeg, roadSeg;
  int status[kNumStatus] = {};
  PinnedBuf<int> hostRoads;  // staging of itsca_download_roads

  // itsca_run_noalloc: the device address of `scratch` on this device, and
  // fixed host staging for its small inputs and read-backs
  unsigned char* scratchPtr = nullptr;
  int naEvOff[8] = {};
  int naZero[2] = {};
  float naVertex[3] = {};
  Seg naClSeg[kLayers] = {};
  Seg naRoadSeg = {};
  int naStatus[kNumStatus] = {};

  // device buffers
  DBuf<float> rx, ry, rz, itemVertex;
  DBuf<int> rid, rmc, evLayerOff, dItemEvent, dRecItem;
  DBuf<Seg> dClSeg, dTrkSeg, dCelSeg, dEdgSeg, dRoadSeg;
  DBuf<float> cz, cphi, cr, sPhi, sR;
  DBuf<float4> cxyrz;
  DBuf<int> cid, cmc, cbin, sBin, sInv, trkCnt, trkOff, trkKept, table;
  DBuf<float4> crzp, crzpS, segGeo;
  DBuf<uint4> subTab;
  DBuf<int> cellCnt, cellOff, trkCand, trkKeep, cellCand;
  DBuf<int2> tFS;
  DBuf<float2> tTP;
  DBuf<int> succCnt, succCand, eFrom, eTo, eTmp;
  DBuf<int> c1, c2, c3, ct1, ct2, clevel, edgeCnt, edgeOff, roadOff;
  DBuf<float4> cnc;
  DBuf<uint32_t> cpaths;
  DBuf<int> pred, rCells, rLabel, rFake;
  DBuf<int> chunkStart, chunkSum, chunkOff, planScratch, dStatus;
  DBuf<int4> chunkInfo;

  void dropGraph()
  {
    if (graphExec) (void)hipGraphExecDestroy(graphExec);
    if (graph) (void)hipGraphDestroy(graph);
    graphExec = nullptr;
    graph = nullptr;
    graphValid = false;
  }

  ~itsca_ctx()
  {
    dropGraph();
    if (timerReady)
      for (auto& e : timer.ev) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  itsca_status fail(int code, const std::string& msg)
  {
    err = msg;
    return code;
  }

  // A road stage that cannot be completed by growing buffers; sets err.
  bool roadFailure()
  {
    if (status[kStatusRootRoads] != 0) {
      err = "a root cell has more than " + std::to_string(kMaxRootRoads) + " roads";
      return true;
    }
    if (status[kStatusUndecodable] != 0) {
      err = "a road could not be decoded from the path counts";
      return true;
    }
    return false;
  }

  void estimateCapacities()
  {
    double t = 0, c = 0;
    for (int w = 0; w < W; ++w) {
      double tw = 0;
      for (int l = 0; l < kTrackletLayers; ++l) {
        const double a = clSeg[w * kLayers + l].count, b = clSeg[w * kLayers + l + 1].count;
        tw += std::max({4096.0, 24.0 * std::max(a, b), 3e-3 * a * b});
      }
      t += tw;
      c += std::max(4096.0, 0.25 * tw);
    }
    const double sc = capacityScale > 0 ? capacityScale : 1.0;
    auto clampCap = [](double v) { return (int)std::min(v, kMaxCapacity); };
    tCap = clampCap(t * sc);
    cCap = clampCap(c * sc);
    eCap = cCap;
    rCap = clampCap(std::max(4096.0, 0.5 * c * sc));
  }

  void allocateOutputs()
  {
    dropGraph();  // kernel arguments hold the buffer addresses
    const size_t SC = (size_t)std::max<long>(sortedClusters, 1);
    maxSegs = kLayers * std::max(W, 1);
    maxChunks = (int)((std::max<size_t>({SC, (size_t)tCap, (size_t)cCap}) + kMinChunk - 1) / kMinChunk) + 8 * W + 1;
    HIP_TRY(cxyrz.reserve(SC));
    HIP_TRY(cz.reserve(SC));
    HIP_TRY(cphi.reserve(SC));
    HIP_TRY(cr.reserve(SC));
    HIP_TRY(sPhi.reserve(SC));
    HIP_TRY(sR.reserve(SC));
    HIP_TRY(cid.reserve(SC));
    HIP_TRY(cmc.reserve(SC));
    HIP_TRY(cbin.reserve(SC));
    HIP_TRY(sBin.reserve(SC));
    HIP_TRY(sInv.reserve(SC));
    HIP_TRY(trkCnt.reserve(SC));
    HIP_TRY(trkOff.reserve(SC));
    HIP_TRY(trkKept.reserve(SC));
    HIP_TRY(crzp.reserve(SC));
    HIP_TRY(crzpS.reserve(SC));
    HIP_TRY(subTab.reserve((size_t)std::max(W, 1) * kLayers * kBins));
    HIP_TRY(trkCand.reserve(SC * kTrkSlots));
    HIP_TRY(trkKeep.reserve(SC * kTrkSlots));
    HIP_TRY(table.reserve((size_t)W * kLayers * kTableSize));
    HIP_TRY(segGeo.reserve((size_t)std::max(W, 1) * kLayers));
    HIP_TRY(tFS.reserve(tCap));
    HIP_TRY(tTP.reserve((size_t)tCap + kTrackletPad));
    HIP_TRY(cellCnt.reserve(tCap));
    HIP_TRY(cellOff.reserve(tCap));
    HIP_TRY(cellCand.reserve((size_t)tCap * kCellSlots));
    HIP_TRY(c1.reserve(cCap));
    HIP_TRY(c2.reserve(cCap));
    HIP_TRY(c3.reserve(cCap));
    HIP_TRY(ct1.reserve(cCap));
    HIP_TRY(ct2.reserve(cCap));
    HIP_TRY(clevel.reserve(cCap));
    HIP_TRY(cnc.reserve(cCap));
    HIP_TRY(cpaths.reserve(cCap));
    HIP_TRY(edgeCnt.reserve(cCap));
    HIP_TRY(edgeOff.reserve(cCap));
    HIP_TRY(roadOff.reserve(cCap));
    HIP_TRY(pred.reserve(eCap));
    HIP_TRY(succCnt.reserve(cCap));
    HIP_TRY(succCand.reserve((size_t)cCap * kEdgeSlots));
    HIP_TRY(eFrom.reserve(eCap));
    HIP_TRY(eTo.reserve(eCap));
    HIP_TRY(eTmp.reserve(eCap));
    HIP_TRY(rCells.reserve((size_t)rCap * kCellLayers));
    HIP_TRY(rLabel.reserve(rCap));
    HIP_TRY(rFake.reserve(rCap));
    HIP_TRY(dTrkSeg.reserve((size_t)W * kTrackletLayers));
    HIP_TRY(dCelSeg.reserve((size_t)W * kCellLayers));
    HIP_TRY(dEdgSeg.reserve((size_t)W * (kCellLayers - 1)));
    HIP_TRY(dRoadSeg.reserve((size_t)std::max(W, 1)));
    HIP_TRY(chunkStart.reserve((size_t)kNumStages * (maxSegs + 1)));
    HIP_TRY(chunkSum.reserve((size_t)kNumStages * maxChunks));
    HIP_TRY(chunkOff.reserve((size_t)kNumStages * maxChunks));
    HIP_TRY(chunkInfo.reserve((size_t)kNumStages * maxChunks));
    HIP_TRY(planScratch.reserve(maxSegs));
    HIP_TRY(dStatus.reserve(kNumStatus));
  }

  void ensureMlp()
  {
    if (mlpMaster.p) return;
    HIP_TRY(mlpMaster.reserve(kMlpParams));
    HIP_TRY(mlpMom1.reserve(kMlpParams));
    HIP_TRY(mlpMom2.reserve(kMlpParams));
    HIP_TRY(mlpWq.reserve(kMlpParams));
    HIP_TRY(mlpWbf.reserve(kMlpParams));
    HIP_TRY(mlpGrad.reserve(kMlpParams));
    HIP_TRY(mlpLoss.reserve(1));
    HIP_TRY(mlpPartial.reserve((size_t)256 * (kMlpParams + 2)));
    std::vector<float> zero(kMlpParams, 0.0f);
    std::vector<uint16_t> zero16(kMlpParams, 0);
    HIP_TRY(hipMemcpyAsync(mlpMaster.p, zero.data(), kMlpParams * 4, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(mlpMom1.p, zero.data(), kMlpParams * 4, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(mlpMom2.p, zero.data(), kMlpParams * 4, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(mlpWq.p, zero.data(), kMlpParams * 4, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(mlpWbf.p, zero16.data(), kMlpParams * 2, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
  }

  MlpBuffers mlpBuffers() const
  {
    return MlpBuffers{mlpMaster.p, mlpMom1.p, mlpMom2.p, mlpWq.p, mlpWbf.p, mlpPartial.p, 256};
  }

  void setParams(const float* p)
  {
    ensureMlp();
    std::vector<float> q(kMlpParams);
    std::vector<uint16_t> b(kMlpParams);
    for (int i = 0; i < kMlpParams; ++i) {
      q[i] = roundBf16(p[i]);
      b[i] = toBf16(p[i]);
    }
    std::vector<float> zero(kMlpParams, 0.0f);
    HIP_TRY(hipMemcpyAsync(mlpMaster.p, p, kMlpParams * 4, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(mlpWq.p, q.data(), kMlpParams * 4, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(mlpWbf.p, b.data(), kMlpParams * 2, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(mlpMom1.p, zero.data(), kMlpParams * 4, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(mlpMom2.p, zero.data(), kMlpParams * 4, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    mlpStep = 0;
    dropGraph();
  }

  // the learned form's phi gates (filtered runs only)
  void setGates(DevState& s) const
  {
    s.gated = 0;
    for (int l = 0; l < kTrackletLayers; ++l) {
      s.phiGate[l] = phiGate[l];
      if (filterOn && phiGate[l] < kPhiCut) s.gated = 1;
    }
  }

  DevState state() const
  {
    DevState s{};
    s.W = W;
    s.maxSegs = maxSegs;
    s.chunk = blockSize;
    s.itemEvent = dItemEvent.p;
    s.recItem = dRecItem.p;
    s.itemVertex = itemVertex.p;
    s.rx = rx.p;
    s.ry = ry.p;
    s.rz = rz.p;
    s.rid = rid.p;
    s.rmc = rmc.p;
    s.evLayerOff = evLayerOff.p;
    s.clSeg = dClSeg.p;
    s.cxyrz = cxyrz.p;
    s.cz = cz.p;
    s.cphi = cphi.p;
    s.cr = cr.p;
    s.crzp = crzp.p;
    s.crzpS = crzpS.p;
    s.subTab = subTab.p;
    s.zSub = zSub;
    s.phiSub = phiSub;
    s.segGeo = segGeo.p;
    s.cid = cid.p;
    s.cmc = cmc.p;
    s.cbin = cbin.p;
    s.sPhi = sPhi.p;
    s.sR = sR.p;
    s.sBin = sBin.p;
    s.sInv = sInv.p;
    s.table = table.p;
    s.trkCnt = trkCnt.p;
    s.trkOff = trkOff.p;
    s.trkKept = trkKept.p;
    s.mlpW = mlpWbf.p;
    setGates(s);
    s.trkCand = trkCand.p;
    s.trkKeep = trkKeep.p;
    s.trkSeg = dTrkSeg.p;
    s.tFS = tFS.p;
    s.tTP = tTP.p;
    s.tCap = tCap;
    s.cellCnt = cellCnt.p;
    s.cellOff = cellOff.p;
    s.cellCand = cellCand.p;
    s.celSeg = dCelSeg.p;
    s.c1 = c1.p;
    s.c2 = c2.p;
    s.c3 = c3.p;
    s.ct1 = ct1.p;
    s.ct2 = ct2.p;
    s.clevel = clevel.p;
    s.cnc = cnc.p;
    s.cpaths = cpaths.p;
    s.cCap = cCap;
    s.succCnt = succCnt.p;
    s.cFirst = succCnt.p;
    s.succCand = succCand.p;
    s.eFrom = eFrom.p;
    s.eTo = eTo.p;
    s.eTmp = eTmp.p;
    s.edgeCnt = edgeCnt.p;
    s.edgeOff = edgeOff.p;
    s.edgSeg = dEdgSeg.p;
    s.pred = pred.p;
    s.eCap = eCap;
    s.roadOff = roadOff.p;
    s.roadSeg = dRoadSeg.p;
    s.rCells = rCells.p;
    s.rLabel = rLabel.p;
    s.rFake = rFake.p;
    s.rCap = rCap;
    s.chunkStart = chunkStart.p;
    s.chunkSum = chunkSum.p;
    s.chunkOff = chunkOff.p;
    s.chunkInfo = chunkInfo.p;
    s.maxChunks = maxChunks;
    s.status = dStatus.p;
    return s;
  }

  void enqueue()
  {
    int maxN = 0;
    for (const Seg& g : clSeg) maxN = std::max(maxN, g.count);
    LaunchConfig cfg{persistentBlocks, planScratch.p, maxN, filterOn, threshold, filterFused, cus};
    if (profiling && !timerReady) {
      for (auto& e : timer.ev) HIP_TRY(hipEventCreate(&e));
      timerReady = true;
    }
    if (useGraph && !profiling) {
      if (!graphValid) {
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        launchPipeline(state(), cfg, stream, nullptr);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(stream, &g);
        HIP_TRY(e);
        graph = g;
        HIP_TRY(hipGraphInstantiate(&graphExec, graph, nullptr, nullptr, 0));
        graphValid = true;
      }
      HIP_TRY(hipGraphLaunch(graphExec, stream));
    } else {
      launchPipeline(state(), cfg, stream, profiling ? &timer : nullptr);
    }
    HIP_TRY(hipGetLastError());
  }

  void readBack()
  {
    trkSeg.resize((size_t)W * kTrackletLayers);
    celSeg.resize((size_t)W * kCellLayers);
    edgSeg.resize((size_t)W * (kCellLayers - 1));
    roadSeg.resize(W);
    HIP_TRY(hipMemcpyAsync(trkSeg.data(), dTrkSeg.p, trkSeg.size() * sizeof(Seg), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(celSeg.data(), dCelSeg.p, celSeg.size() * sizeof(Seg), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(edgSeg.data(), dEdgSeg.p, edgSeg.size() * sizeof(Seg), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(roadSeg.data(), dRoadSeg.p, roadSeg.size() * sizeof(Seg), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(status, dStatus.p, sizeof status, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
  }
};

namespace {

template <typename F>
itsca_status guarded(itsca_ctx* ctx, const char* what, F&& body)
{
  if (!ctx) return ITSCA_ERR_ARGUMENT;
  try {
    return body();
  } catch (const HipError& e) {
    return ctx->fail(ITSCA_ERR_HIP, std::string(what) + ": " + e.where + ": " + hipGetErrorString(e.code));
  } catch (const std::bad_alloc&) {
    return ctx->fail(ITSCA_ERR_ARGUMENT, std::string(what) + ": host allocation failed");
  }
}

}  // namespace

extern "C" {

itsca_status itsca_create(itsca_ctx** out, const itsca_config* cfg)
{
  if (!out) return ITSCA_ERR_ARGUMENT;
  *out = nullptr;
  itsca_ctx* ctx = new itsca_ctx();
  if (cfg) {
    ctx->device = cfg->device;
    ctx->persistentBlocks = cfg->persistent_blocks;
    ctx->capacityScale = cfg->capacity_scale;
  }
  const itsca_status st = guarded(ctx, "itsca_create", [&]() -> itsca_status {
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (ctx->device < 0 || ctx->device >= n)
      return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_create: no HIP device " + std::to_string(ctx->device));
    HIP_TRY(hipSetDevice(ctx->device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, ctx->device));
    ctx->cus = prop.multiProcessorCount;
    if (ctx->persistentBlocks <= 0) ctx->persistentBlocks = 8 * ctx->cus;
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->scratchPtr = static_cast<unsigned char*>(scratchAddress(ctx->device));
    return ITSCA_OK;
  });
  // On failure the context is still returned so that the caller can read
  // itsca_last_error; it must be destroyed like any other.
  *out = ctx;
  return st;
}

void itsca_destroy(itsca_ctx* ctx)
{
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  delete ctx;
}

const char* itsca_last_error(const itsca_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

itsca_status itsca_clear_events(itsca_ctx* ctx)
{
  if (!ctx) return ITSCA_ERR_ARGUMENT;
  ctx->hx.clear();
  ctx->hy.clear();
  ctx->hz.clear();
  ctx->hid.clear();
  ctx->hmc.clear();
  ctx->hvert.clear();
  ctx->hLayerOff.clear();
  ctx->hEvVert.clear();
  ctx->E = 0;
  ctx->ran = false;
  return ITSCA_OK;
}

itsca_status itsca_add_event(itsca_ctx* ctx, const float* x, const float* y, const float* z,
                             const int32_t* cluster_id, const int32_t* mc_id, const int32_t* layer_offsets,
                             const float* vertices, int num_vertices, int* event_index)
{
  if (!ctx) return ITSCA_ERR_ARGUMENT;
  if (!layer_offsets || num_vertices < 0 || (num_vertices > 0 && !vertices))
    return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_add_event: missing layer offsets or vertices");
  const int total = layer_offsets[kLayers] - layer_offsets[0];
  for (int l = 0; l < kLayers; ++l)
    if (layer_offsets[l + 1] < layer_offsets[l])
      return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_add_event: layer offsets must be non-decreasing");
  if (total > 0 && (!x || !y || !z || !cluster_id || !mc_id))
    return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_add_event: missing cluster arrays");
  const int base = (int)ctx->hx.size();
  if ((long)base + total > 0x7fffffffL) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_add_event: batch too large");
  const int o = layer_offsets[0];
  for (int v = 0; v < 3 * num_vertices; ++v)
    if (!std::isfinite(vertices[v]))
      return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_add_event: vertex " + std::to_string(v / 3) +
                                               " has a non-finite coordinate");
  const size_t want = (size_t)base + (size_t)total;
  hipError_t e = ctx->hx.reserve(want);
  if (e == hipSuccess) e = ctx->hy.reserve(want);
  if (e == hipSuccess) e = ctx->hz.reserve(want);
  if (e == hipSuccess) e = ctx->hid.reserve(want);
  if (e == hipSuccess) e = ctx->hmc.reserve(want);
  if (e != hipSuccess)
    return ctx->fail(ITSCA_ERR_HIP, std::string("itsca_add_event: pinned host staging: ") + hipGetErrorString(e));
  // The coordinates are copied and checked in one pass and committed only
  // when all are finite.  Non-finite coordinates have no defined bin: the
  // reference's conversions of NaN are undefined behaviour in C++, so they
  // are refused here.
  float* dx = ctx->hx.data() + base;
  float* dy = ctx->hy.data() + base;
  float* dz = ctx->hz.data() + base;
  bool finite = true;
  for (int i = 0; i < total; ++i) {
    const float a = x[o + i], b = y[o + i], c = z[o + i];
    dx[i] = a;
    dy[i] = b;
    dz[i] = c;
    finite = finite & std::isfinite(a) & std::isfinite(b) & std::isfinite(c);
  }
  if (!finite)
    for (int i = 0; i < total; ++i)
      if (!std::isfinite(dx[i]) || !std::isfinite(dy[i]) || !std::isfinite(dz[i]))
        return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_add_event: cluster " + std::to_string(i) +
                                                 " has a non-finite coordinate");
  if (total > 0) {
    std::memcpy(ctx->hid.data() + base, cluster_id + o, (size_t)total * sizeof(int));
    std::memcpy(ctx->hmc.data() + base, mc_id + o, (size_t)total * sizeof(int));
  }
  ctx->hx.n = ctx->hy.n = ctx->hz.n = ctx->hid.n = ctx->hmc.n = want;
  for (int l = 0; l < 8; ++l) ctx->hLayerOff.push_back(base + layer_offsets[l] - o);
  ctx->hEvVert.push_back((int)ctx->hvert.size() / 3);
  ctx->hvert.insert(ctx->hvert.end(), vertices, vertices + 3 * num_vertices);
  if (event_index) *event_index = ctx->E;
  ++ctx->E;
  ctx->ran = false;
  return ITSCA_OK;
}

itsca_status itsca_upload(itsca_ctx* ctx)
{
  return guarded(ctx, "itsca_upload", [&]() -> itsca_status {
    HIP_TRY(hipSetDevice(ctx->device));
    // Work items: every vertex of every event, in event order.  A vertex
    // with the same x and y (bit for bit) as an earlier vertex of its event
    // reads that vertex's prepared clusters: phi and r are taken about
    // (x, y) and the bins use the absolute z, so they are identical.
    ctx->itemEvent.clear();
    ctx->recItem.clear();
    ctx->clSeg.clear();
    std::vector<float> itemVtx;
    long sc = 0;
    const int nv = (int)ctx->hvert.size() / 3;
    for (int e = 0; e < ctx->E; ++e) {
      const int v0 = ctx->hEvVert[e];
      const int v1 = e + 1 < ctx->E ? ctx->hEvVert[e + 1] : nv;
      const int firstItem = (int)ctx->itemEvent.size();
      for (int v = v0; v < v1; ++v) {
        int rec = (int)ctx->itemEvent.size();
        for (int u = v0; u < v; ++u)
          if (std::memcmp(&ctx->hvert[3 * u], &ctx->hvert[3 * v], 2 * sizeof(float)) == 0) {
            rec = ctx->recItem[firstItem + (u - v0)];
            break;
          }
        ctx->recItem.push_back(rec);
        ctx->itemEvent.push_back(e);
        itemVtx.insert(itemVtx.end(), {ctx->hvert[3 * v], ctx->hvert[3 * v + 1], ctx->hvert[3 * v + 2]});
        for (int l = 0; l < kLayers; ++l) {
          const int n = ctx->hLayerOff[8 * e + l + 1] - ctx->hLayerOff[8 * e + l];
          ctx->clSeg.push_back(Seg{(int)sc, n});
          sc += n;
        }
      }
    }
    if (sc > 0x7fffffffL) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_upload: batch too large");
    ctx->W = (int)ctx->itemEvent.size();
    ctx->sortedClusters = sc;
    const size_t NR = std::max<size_t>(ctx->hx.size(), 1);
    HIP_TRY(ctx->rx.reserve(NR));
    HIP_TRY(ctx->ry.reserve(NR));
    HIP_TRY(ctx->rz.reserve(NR));
    HIP_TRY(ctx->rid.reserve(NR));
    HIP_TRY(ctx->rmc.reserve(NR));
    HIP_TRY(ctx->evLayerOff.reserve(std::max<size_t>(ctx->hLayerOff.size(), 1)));
    HIP_TRY(ctx->dItemEvent.reserve(std::max(ctx->W, 1)));
    HIP_TRY(ctx->dRecItem.reserve(std::max(ctx->W, 1)));
    HIP_TRY(ctx->itemVertex.reserve(std::max<size_t>(itemVtx.size(), 1)));
    HIP_TRY(ctx->dClSeg.reserve(std::max<size_t>(ctx->clSeg.size(), 1)));
    auto put = [&](void* dst, const void* src, size_t bytes) {
      if (bytes) HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    };
    put(ctx->rx.p, ctx->hx.data(), ctx->hx.size() * 4);
    put(ctx->ry.p, ctx->hy.data(), ctx->hy.size() * 4);
    put(ctx->rz.p, ctx->hz.data(), ctx->hz.size() * 4);
    put(ctx->rid.p, ctx->hid.data(), ctx->hid.size() * 4);
    put(ctx->rmc.p, ctx->hmc.data(), ctx->hmc.size() * 4);
    put(ctx->evLayerOff.p, ctx->hLayerOff.data(), ctx->hLayerOff.size() * 4);
    put(ctx->dItemEvent.p, ctx->itemEvent.data(), ctx->itemEvent.size() * 4);
    put(ctx->dRecItem.p, ctx->recItem.data(), ctx->recItem.size() * 4);
    put(ctx->itemVertex.p, itemVtx.data(), itemVtx.size() * 4);
    put(ctx->dClSeg.p, ctx->clSeg.data(), ctx->clSeg.size() * sizeof(Seg));
    ctx->estimateCapacities();
    ctx->allocateOutputs();
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->ran = false;
    return ITSCA_OK;
  });
}

itsca_status itsca_run_async(itsca_ctx* ctx)
{
  return guarded(ctx, "itsca_run_async", [&]() -> itsca_status {
    if (ctx->W == 0) return ctx->fail(ITSCA_ERR_STATE, "itsca_run_async: nothing uploaded");
    HIP_TRY(hipSetDevice(ctx->device));
    ctx->enqueue();
    return ITSCA_OK;
  });
}

itsca_status itsca_finish(itsca_ctx* ctx)
{
  return guarded(ctx, "itsca_finish", [&]() -> itsca_status {
    HIP_TRY(hipSetDevice(ctx->device));
    ctx->readBack();
    if (ctx->roadFailure()) return ctx->fail(ITSCA_ERR_CAPACITY, "itsca_finish: " + ctx->err);
    for (int k = 0; k < kNumStatus; ++k)
      if (ctx->status[k] != 0)
        return ctx->fail(ITSCA_ERR_CAPACITY, "itsca_finish: output capacity exceeded (stage " + std::to_string(k) +
                                                 ", needed " + std::to_string(ctx->status[k]) + ")");
    ctx->ran = true;
    return ITSCA_OK;
  });
}

itsca_status itsca_run(itsca_ctx* ctx)
{
  return guarded(ctx, "itsca_run", [&]() -> itsca_status {
    if (ctx->W == 0) return ctx->fail(ITSCA_ERR_STATE, "itsca_run: nothing uploaded");
    HIP_TRY(hipSetDevice(ctx->device));
    for (int attempt = 0; attempt < 6; ++attempt) {
      ctx->enqueue();
      ctx->readBack();
      if (ctx->roadFailure()) return ctx->fail(ITSCA_ERR_CAPACITY, "itsca_run: " + ctx->err);
      bool grow = false;
      bool tooLarge = false;
      auto need = [&](int slot, int& cap) {
        if (ctx->status[slot] != 0) {
          // the scan reports at most INT32_MAX outputs: beyond the largest
          // capacity the run cannot succeed
          if ((double)ctx->status[slot] > kMaxCapacity) tooLarge = true;
          cap = (int)std::min(kMaxCapacity, 1.25 * (double)ctx->status[slot] + 4096);
          grow = true;
        }
      };
      need(kStatusOverflowTracklets, ctx->tCap);
      need(kStatusOverflowCells, ctx->cCap);
      need(kStatusOverflowEdges, ctx->eCap);
      need(kStatusOverflowRoads, ctx->rCap);
      if (tooLarge)
        return ctx->fail(ITSCA_ERR_CAPACITY, "itsca_run: a stage needs more than " +
                                                 std::to_string((long long)kMaxCapacity) + " outputs");
      if (!grow) {
        ctx->ran = true;
        return ITSCA_OK;
      }
      ctx->allocateOutputs();
    }
    return ctx->fail(ITSCA_ERR_CAPACITY, "itsca_run: outputs kept exceeding the grown capacity");
  });
}

int itsca_num_items(const itsca_ctx* ctx) { return ctx ? ctx->W : 0; }

itsca_status itsca_download_roads(itsca_ctx* ctx, int32_t* rows, size_t cap_rows, int32_t* first_road,
                                  size_t* total)
{
  return guarded(ctx, "itsca_download_roads", [&]() -> itsca_status {
    if (!ctx->ran) return ctx->fail(ITSCA_ERR_STATE, "itsca_download_roads: no completed run");
    if (!first_road || !total) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_download_roads: missing output");
    // roads are compact in item order (one scan over the batch), so one copy
    // per array covers every item
    size_t n = 0;
    for (int w = 0; w < ctx->W; ++w) {
      first_road[w] = (int32_t)n;
      n += (size_t)ctx->roadSeg[w].count;
    }
    first_road[ctx->W] = (int32_t)n;
    *total = n;
    if (n > cap_rows) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_download_roads: destination too small");
    if (n == 0) return ITSCA_OK;
    if (!rows) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_download_roads: missing rows");
    const size_t base = (size_t)ctx->roadSeg[0].base;
    for (int w = 1; w < ctx->W; ++w)
      if ((size_t)ctx->roadSeg[w].base != base + (size_t)first_road[w])
        return ctx->fail(ITSCA_ERR_STATE, "itsca_download_roads: roads are not in item order");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(ctx->hostRoads.reserve(n * (kCellLayers + 2)));
    ctx->hostRoads.n = n * (kCellLayers + 2);
    int* cells = ctx->hostRoads.data();
    int* lab = cells + n * kCellLayers;
    int* fake = lab + n;
    HIP_TRY(hipMemcpyAsync(cells, ctx->rCells.p + base * kCellLayers, n * kCellLayers * 4, hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipMemcpyAsync(lab, ctx->rLabel.p + base, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(fake, ctx->rFake.p + base, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < n; ++i) {
      int32_t* r = rows + i * 8;
      int size = 0;
      for (int k = 0; k < kCellLayers; ++k) {
        r[k] = cells[i * kCellLayers + k];
        size += r[k] >= 0 ? 1 : 0;
      }
      r[5] = size;
      r[6] = lab[i];
      r[7] = fake[i];
    }
    return ITSCA_OK;
  });
}

itsca_status itsca_get_capacities(itsca_ctx* ctx, itsca_capacities* out)
{
  if (!ctx || !out) return ITSCA_ERR_ARGUMENT;
  out->tracklets = ctx->tCap;
  out->cells = ctx->cCap;
  out->neighbours = ctx->eCap;
  out->roads = ctx->rCap;
  return ITSCA_OK;
}

namespace {

// Lays out one work item's intermediate arrays in the scratch array of
// `bytes` bytes at device address `base`, for n sorted clusters: the
// per-cluster arrays first, then the rest split between tracklets, cells,
// edges and roads in the proportions 10 : 1 : 1 : 1 (a central event has
// about 9 tracklets, 1 edge and 0.2 roads per cell).  Returns false if the
// clusters alone do not leave room for them.
// keptLists: the filter is on, so the chunks' kept lists (trkKeep) need room.
bool carveNoAlloc(unsigned char* base, size_t bytes, int n, bool keptLists, DevState& s, int*& planScratch)
{
  size_t off = 0;
  bool fits = true;
  auto take = [&](size_t count, size_t elem) -> void* {
    off = (off + 255) & ~(size_t)255;
    void* p = base + off;
    off += count * elem;
    if (off > bytes) fits = false;
    return p;
  };
  const size_t N = (size_t)std::max(n, 1);
  s.itemEvent = static_cast<int*>(take(1, 4));
  s.recItem = static_cast<int*>(take(1, 4));
  s.itemVertex = static_cast<float*>(take(3, 4));
  s.evLayerOff = static_cast<int*>(take(8, 4));
  s.clSeg = static_cast<Seg*>(take(kLayers, sizeof(Seg)));
  s.trkSeg = static_cast<Seg*>(take(kTrackletLayers, sizeof(Seg)));
  s.celSeg = static_cast<Seg*>(take(kCellLayers, sizeof(Seg)));
  s.edgSeg = static_cast<Seg*>(take(kCellLayers - 1, sizeof(Seg)));
  s.roadSeg = static_cast<Seg*>(take(1, sizeof(Seg)));
  s.status = static_cast<int*>(take(kNumStatus, 4));
  planScratch = static_cast<int*>(take(kLayers, 4));
  s.table = static_cast<int*>(take((size_t)kLayers * kTableSize, 4));
  s.subTab = static_cast<uint4*>(take((size_t)kLayers * kBins, sizeof(uint4)));
  s.segGeo = static_cast<float4*>(take(kLayers, sizeof(float4)));
  s.cxyrz = static_cast<float4*>(take(N, sizeof(float4)));
  s.crzp = static_cast<float4*>(take(N, sizeof(float4)));
  s.crzpS = static_cast<float4*>(take(N, sizeof(float4)));
  s.cz = static_cast<float*>(take(N, 4));
  s.cphi = static_cast<float*>(take(N, 4));
  s.cr = static_cast<float*>(take(N, 4));
  s.sPhi = static_cast<float*>(take(N, 4));
  s.sR = static_cast<float*>(take(N, 4));
  s.cid = static_cast<int*>(take(N, 4));
  s.cmc = static_cast<int*>(take(N, 4));
  s.cbin = static_cast<int*>(take(N, 4));
  s.sBin = static_cast<int*>(take(N, 4));
  s.sInv = static_cast<int*>(take(N, 4));
  s.trkCnt = static_cast<int*>(take(N, 4));
  s.trkOff = static_cast<int*>(take(N, 4));
  s.trkKept = static_cast<int*>(take(N, 4));
  s.trkCand = static_cast<int*>(take(N * kTrkSlots, 4));
  s.trkKeep = keptLists ? static_cast<int*>(take(N * kTrkSlots, 4)) : nullptr;
  s.chunkStart = static_cast<int*>(take((size_t)kNumStages * (kLayers + 1), 4));
  if (!fits) return false;
  // bytes per unit k of capacity: 10 tracklets, 1 cell, 1 edge, 1 road, and
  // two chunk-plan ints per stage for every kMinChunk tracklets
  const size_t perTracklet = 8 + 8 + 4 + 4 + 4 * kCellSlots;
  const size_t perCell = 6 * 4 + 16 + 4 + 3 * 4 + 4 + 4 * kEdgeSlots;
  const size_t perEdge = 4 * 4, perRoad = 4 * kCellLayers + 4 + 4;
  const size_t perUnit = 10 * perTracklet + perCell + perEdge + perRoad + 6 * 4 * kNumStages * 10 / kMinChunk + 64;
  const size_t left = bytes - off;
  const size_t k = std::min<size_t>(left / perUnit, (size_t)(kMaxCapacity / 10));
  if (k < 64) return false;
  s.tCap = (int)(10 * k);
  s.cCap = s.eCap = s.rCap = (int)k;
  s.maxSegs = kLayers;
  s.maxChunks = (int)((std::max<size_t>(N, 10 * k) + kMinChunk - 1) / kMinChunk) + 8 + 1;
  s.chunkSum = static_cast<int*>(take((size_t)kNumStages * s.maxChunks, 4));
  s.chunkOff = static_cast<int*>(take((size_t)kNumStages * s.maxChunks, 4));
  s.chunkInfo = static_cast<int4*>(take((size_t)kNumStages * s.maxChunks, sizeof(int4)));
  s.tFS = static_cast<int2*>(take(s.tCap, sizeof(int2)));
  s.tTP = static_cast<float2*>(take((size_t)s.tCap + kTrackletPad, sizeof(float2)));
  s.cellCnt = static_cast<int*>(take(s.tCap, 4));
  s.cellOff = static_cast<int*>(take(s.tCap, 4));
  s.cellCand = static_cast<int*>(take((size_t)s.tCap * kCellSlots, 4));
  s.c1 = static_cast<int*>(take(s.cCap, 4));
  s.c2 = static_cast<int*>(take(s.cCap, 4));
  s.c3 = static_cast<int*>(take(s.cCap, 4));
  s.ct1 = static_cast<int*>(take(s.cCap, 4));
  s.ct2 = static_cast<int*>(take(s.cCap, 4));
  s.clevel = static_cast<int*>(take(s.cCap, 4));
  s.cnc = static_cast<float4*>(take(s.cCap, sizeof(float4)));
  s.cpaths = static_cast<uint32_t*>(take(s.cCap, 4));
  s.edgeCnt = static_cast<int*>(take(s.cCap, 4));
  s.edgeOff = static_cast<int*>(take(s.cCap, 4));
  s.roadOff = static_cast<int*>(take(s.cCap, 4));
  s.succCnt = static_cast<int*>(take(s.cCap, 4));
  s.cFirst = s.succCnt;
  s.succCand = static_cast<int*>(take((size_t)s.cCap * kEdgeSlots, 4));
  s.pred = static_cast<int*>(take(s.eCap, 4));
  s.eFrom = static_cast<int*>(take(s.eCap, 4));
  s.eTo = static_cast<int*>(take(s.eCap, 4));
  s.eTmp = static_cast<int*>(take(s.eCap, 4));
  s.rCells = static_cast<int*>(take((size_t)s.rCap * kCellLayers, 4));
  s.rLabel = static_cast<int*>(take(s.rCap, 4));
  s.rFake = static_cast<int*>(take(s.rCap, 4));
  return fits;
}

}  // namespace

itsca_status itsca_run_noalloc(itsca_ctx* ctx, const itsca_device_event* ev, int32_t* roads, int roads_cap,
                               int* num_roads)
{
  if (!ctx) return ITSCA_ERR_ARGUMENT;
  if (!ev || !num_roads || roads_cap < 0 || (roads_cap > 0 && !roads))
    return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_run_noalloc: missing event, road buffer or count");
  *num_roads = 0;
  const int32_t* offs = ev->layer_offsets;
  int maxLayer = 0;
  for (int l = 0; l < kLayers; ++l) {
    if (offs[l + 1] < offs[l] || offs[l] < 0)
      return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_run_noalloc: layer offsets must be non-negative and non-decreasing");
    maxLayer = std::max(maxLayer, offs[l + 1] - offs[l]);
  }
  const long n = (long)offs[kLayers] - offs[0];
  if (n > ITSCA_NOALLOC_MAX_CLUSTERS)
    return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_run_noalloc: " + std::to_string(n) + " clusters, more than the " +
                                             std::to_string(ITSCA_NOALLOC_MAX_CLUSTERS) + " this entry point takes");
  if (n > 0 && (!ev->x || !ev->y || !ev->z || !ev->cluster_id || !ev->mc_id))
    return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_run_noalloc: missing cluster arrays");
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(ev->vertex[k]))
      return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_run_noalloc: the vertex has a non-finite coordinate");
  return guarded(ctx, "itsca_run_noalloc", [&]() -> itsca_status {
    HIP_TRY(hipSetDevice(ctx->device));
    if (!ctx->scratchPtr) return ctx->fail(ITSCA_ERR_STATE, "itsca_run_noalloc: no device scratch");
    DevState s{};
    int* planScratch = nullptr;
    const bool filtered = ctx->filterOn && ctx->mlpWbf.p != nullptr;
    if (!carveNoAlloc(ctx->scratchPtr, ITSCA_NOALLOC_SCRATCH_BYTES, (int)n, filtered, s, planScratch))
      return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_run_noalloc: the event does not fit the device scratch");
    s.W = 1;
    s.chunk = ctx->blockSize;
    s.zSub = ctx->zSub;
    s.phiSub = ctx->phiSub;
    s.rx = ev->x;
    s.ry = ev->y;
    s.rz = ev->z;
    s.rid = ev->cluster_id;
    s.rmc = ev->mc_id;
    s.mlpW = ctx->mlpWbf.p;
    ctx->setGates(s);
    // small inputs: one work item, its layer segments, the vertex
    for (int l = 0; l < 8; ++l) ctx->naEvOff[l] = offs[l];
    for (int l = 0; l < kLayers; ++l) ctx->naClSeg[l] = Seg{offs[l] - offs[0], offs[l + 1] - offs[l]};
    for (int k = 0; k < 3; ++k) ctx->naVertex[k] = ev->vertex[k];
    ctx->naZero[0] = ctx->naZero[1] = 0;
    HIP_TRY(hipMemcpyAsync(const_cast<int*>(s.evLayerOff), ctx->naEvOff, sizeof ctx->naEvOff, hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(hipMemcpyAsync(const_cast<Seg*>(s.clSeg), ctx->naClSeg, sizeof ctx->naClSeg, hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(hipMemcpyAsync(const_cast<float*>(s.itemVertex), ctx->naVertex, sizeof ctx->naVertex,
                           hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(const_cast<int*>(s.itemEvent), ctx->naZero, 4, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(const_cast<int*>(s.recItem), ctx->naZero, 4, hipMemcpyHostToDevice, ctx->stream));
    LaunchConfig cfg{ctx->persistentBlocks, planScratch, maxLayer, ctx->filterOn && ctx->mlpWbf.p != nullptr,
                     ctx->threshold, ctx->filterFused, ctx->cus};
    launchPipeline(s, cfg, ctx->stream, nullptr);
    if (roads_cap > 0) launchRoadsOut(s, 0, roads, roads_cap, ctx->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(&ctx->naRoadSeg, s.roadSeg, sizeof(Seg), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->naStatus, s.status, sizeof ctx->naStatus, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const int* st = ctx->naStatus;
    if (st[kStatusRootRoads] != 0)
      return ctx->fail(ITSCA_ERR_CAPACITY,
                       "itsca_run_noalloc: a root cell has more than " + std::to_string(kMaxRootRoads) + " roads");
    if (st[kStatusUndecodable] != 0)
      return ctx->fail(ITSCA_ERR_CAPACITY, "itsca_run_noalloc: a road could not be decoded from the path counts");
    for (int k = 0; k < kNumStatus; ++k)
      if (st[k] != 0)
        return ctx->fail(ITSCA_ERR_CAPACITY, "itsca_run_noalloc: stage " + std::to_string(k) + " needs " +
                                                 std::to_string(st[k]) + " outputs, more than the device scratch holds");
    *num_roads = ctx->naRoadSeg.count;
    if (ctx->naRoadSeg.count > roads_cap)
      return ctx->fail(ITSCA_ERR_CAPACITY, "itsca_run_noalloc: " + std::to_string(ctx->naRoadSeg.count) +
                                               " roads, more than the " + std::to_string(roads_cap) + " rows given");
    return ITSCA_OK;
  });
}

itsca_status itsca_get_counts(itsca_ctx* ctx, int item, itsca_counts* out)
{
  if (!ctx || !out) return ITSCA_ERR_ARGUMENT;
  if (!ctx->ran) return ctx->fail(ITSCA_ERR_STATE, "itsca_get_counts: no completed run");
  if (item < 0 || item >= ctx->W) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_get_counts: bad item");
  for (int l = 0; l < kLayers; ++l) out->clusters[l] = ctx->clSeg[item * kLayers + l].count;
  for (int l = 0; l < kTrackletLayers; ++l) out->tracklets[l] = ctx->trkSeg[item * kTrackletLayers + l].count;
  for (int l = 0; l < kCellLayers; ++l) out->cells[l] = ctx->celSeg[item * kCellLayers + l].count;
  for (int l = 0; l < kCellLayers - 1; ++l) out->neighbours[l] = ctx->edgSeg[item * (kCellLayers - 1) + l].count;
  out->roads = ctx->roadSeg[item].count;
  return ITSCA_OK;
}

itsca_status itsca_download(itsca_ctx* ctx, int item, int what, int layer, void* dst, size_t cap_bytes,
                            size_t* written)
{
  return guarded(ctx, "itsca_download", [&]() -> itsca_status {
    if (!ctx->ran) return ctx->fail(ITSCA_ERR_STATE, "itsca_download: no completed run");
    if (item < 0 || item >= ctx->W) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_download: bad item");
    HIP_TRY(hipSetDevice(ctx->device));
    auto layerOk = [&](int lo, int hi) { return layer >= lo && layer <= hi; };
    std::vector<int32_t> rows;
    auto fetch = [&](const auto* src, size_t n) {
      using T = std::remove_cv_t<std::remove_pointer_t<decltype(src)>>;
      std::vector<T> h(n);
      if (n) HIP_TRY(hipMemcpyAsync(h.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
      return h;
    };
    auto interleave = [&](std::initializer_list<const std::vector<int32_t>*> cols, size_t n) {
      rows.resize(n * cols.size());
      size_t c = 0;
      for (const auto* col : cols) {
        for (size_t i = 0; i < n; ++i) rows[i * cols.size() + c] = (*col)[i];
        ++c;
      }
    };
    auto asInt = [](const std::vector<float>& f) {
      std::vector<int32_t> v(f.size());
      std::memcpy(v.data(), f.data(), f.size() * 4);
      return v;
    };
    switch (what) {
      case ITSCA_CLUSTERS: {
        if (!layerOk(0, 6)) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_download: bad layer");
        const Seg g = ctx->clSeg[ctx->recItem[item] * kLayers + layer];
        auto id = fetch(ctx->cid.p + g.base, g.count), bin = fetch(ctx->cbin.p + g.base, g.count),
             mc = fetch(ctx->cmc.p + g.base, g.count);
        auto phi = asInt(fetch(ctx->cphi.p + g.base, g.count)), r = asInt(fetch(ctx->cr.p + g.base, g.count)),
             z = asInt(fetch(ctx->cz.p + g.base, g.count));
        auto xy = fetch(ctx->cxyrz.p + g.base, g.count);
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        std::vector<int32_t> x(g.count), y(g.count);
        for (int i = 0; i < g.count; ++i) {
          std::memcpy(&x[i], &xy[i].x, 4);
          std::memcpy(&y[i], &xy[i].y, 4);
        }
        interleave({&id, &phi, &r, &bin, &x, &y, &z, &mc}, g.count);
        break;
      }
      case ITSCA_INDEX_TABLE: {
        if (!layerOk(1, 6)) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_download: bad layer");
        rows = fetch(ctx->table.p + (size_t)(ctx->recItem[item] * kLayers + layer) * kTableSize, kTableSize);
        break;
      }
      case ITSCA_TRACKLETS: {
        if (!layerOk(0, 5)) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_download: bad layer");
        const Seg g = ctx->trkSeg[item * kTrackletLayers + layer];
        auto fs = fetch(ctx->tFS.p + g.base, g.count);
        auto tp = fetch(ctx->tTP.p + g.base, g.count);
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        std::vector<int32_t> a(g.count), b(g.count), t(g.count), p(g.count);
        for (int i = 0; i < g.count; ++i) {
          a[i] = fs[i].x;
          b[i] = fs[i].y;
          std::memcpy(&t[i], &tp[i].x, 4);
          std::memcpy(&p[i], &tp[i].y, 4);
        }
        interleave({&a, &b, &t, &p}, g.count);
        break;
      }
      case ITSCA_TRACKLET_LOOKUP: {
        if (!layerOk(0, 4)) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_download: bad layer");
        const Seg g = ctx->clSeg[item * kLayers + layer + 1];
        auto cnt = fetch(ctx->trkCnt.p + g.base, g.count), off = fetch(ctx->trkOff.p + g.base, g.count);
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        rows.resize(g.count);
        for (int i = 0; i < g.count; ++i) rows[i] = cnt[i] > 0 ? off[i] : -1;
        break;
      }
      case ITSCA_CELLS: {
        if (!layerOk(0, 4)) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_download: bad layer");
        const Seg g = ctx->celSeg[item * kCellLayers + layer];
        auto a = fetch(ctx->c1.p + g.base, g.count), b = fetch(ctx->c2.p + g.base, g.count),
             c = fetch(ctx->c3.p + g.base, g.count), t1 = fetch(ctx->ct1.p + g.base, g.count),
             t2 = fetch(ctx->ct2.p + g.base, g.count), lv = fetch(ctx->clevel.p + g.base, g.count);
        auto nc = fetch(ctx->cnc.p + g.base, g.count);
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        std::vector<int32_t> nx(g.count), ny(g.count), nz(g.count), cv(g.count);
        for (int i = 0; i < g.count; ++i) {
          std::memcpy(&nx[i], &nc[i].x, 4);
          std::memcpy(&ny[i], &nc[i].y, 4);
          std::memcpy(&nz[i], &nc[i].z, 4);
          std::memcpy(&cv[i], &nc[i].w, 4);
        }
        interleave({&a, &b, &c, &t1, &t2, &nx, &ny, &nz, &cv, &lv}, g.count);
        break;
      }
      case ITSCA_CELL_LOOKUP: {
        if (!layerOk(0, 3)) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_download: bad layer");
        // indexed by the tracklets of layer+1; set only where cell layer
        // layer+1 was computed
        const Seg g = ctx->trkSeg[item * kTrackletLayers + layer + 1];
        bool computed = true;
        for (int k = 0; k <= layer + 2; ++k) computed = computed && ctx->trkSeg[item * kTrackletLayers + k].count > 0;
        rows.assign(g.count, -1);
        if (computed) {
          auto cnt = fetch(ctx->cellCnt.p + g.base, g.count), off = fetch(ctx->cellOff.p + g.base, g.count);
          HIP_TRY(hipStreamSynchronize(ctx->stream));
          for (int i = 0; i < g.count; ++i) rows[i] = cnt[i] > 0 ? off[i] : -1;
        }
        break;
      }
      case ITSCA_NEIGHBOUR_OFFSETS: {
        if (!layerOk(0, 3)) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_download: bad layer");
        const Seg g = ctx->celSeg[item * kCellLayers + layer + 1];
        auto off = fetch(ctx->edgeOff.p + g.base, g.count);
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        rows = off;
        rows.push_back(ctx->edgSeg[item * (kCellLayers - 1) + layer].count);
        break;
      }
      case ITSCA_NEIGHBOURS: {
        if (!layerOk(0, 3)) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_download: bad layer");
        const Seg g = ctx->edgSeg[item * (kCellLayers - 1) + layer];
        rows = fetch(ctx->pred.p + g.base, g.count);
        break;
      }
      case ITSCA_ROADS: {
        const Seg g = ctx->roadSeg[item];
        auto cells = fetch(ctx->rCells.p + (size_t)g.base * kCellLayers, (size_t)g.count * kCellLayers);
        auto lab = fetch(ctx->rLabel.p + g.base, g.count), fake = fetch(ctx->rFake.p + g.base, g.count);
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        rows.resize((size_t)g.count * 8);
        for (int i = 0; i < g.count; ++i) {
          int size = 0;
          for (int k = 0; k < kCellLayers; ++k) {
            rows[(size_t)i * 8 + k] = cells[(size_t)i * kCellLayers + k];
            size += cells[(size_t)i * kCellLayers + k] >= 0 ? 1 : 0;
          }
          rows[(size_t)i * 8 + 5] = size;
          rows[(size_t)i * 8 + 6] = lab[i];
          rows[(size_t)i * 8 + 7] = fake[i];
        }
        break;
      }
      default:
        return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_download: unknown record kind");
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const size_t bytes = rows.size() * 4;
    if (written) *written = bytes;
    if (bytes > cap_bytes) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_download: destination too small");
    if (bytes) std::memcpy(dst, rows.data(), bytes);
    return ITSCA_OK;
  });
}

itsca_status itsca_set_tuning(itsca_ctx* ctx, const itsca_tuning* t)
{
  if (!ctx || !t) return ITSCA_ERR_ARGUMENT;
  if (t->block_size != 64 && t->block_size != 128 && t->block_size != 256)
    return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_set_tuning: block_size must be 64, 128 or 256");
  if (t->persistent_blocks < 0) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_set_tuning: negative persistent_blocks");
  if (t->z_subbins < 0 || t->z_subbins > kMaxZSub)
    return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_set_tuning: z_subbins must be 0..8");
  if (t->phi_subrows != 0 && t->phi_subrows != 1 && t->phi_subrows != 2 && t->phi_subrows != kMaxPhiSub)
    return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_set_tuning: phi_subrows must be 0, 1, 2 or 4");
  const int zSub = t->z_subbins > 0 ? t->z_subbins : kDefaultZSub, phiSub = t->phi_subrows > 0 ? t->phi_subrows : 1;
  if (zSub * phiSub > kMaxSubBins)
    return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_set_tuning: z_subbins * phi_subrows must be at most 15");
  ctx->blockSize = t->block_size;
  ctx->persistentBlocks = t->persistent_blocks > 0 ? t->persistent_blocks : 8 * ctx->cus;
  ctx->zSub = zSub;
  ctx->phiSub = phiSub;
  ctx->dropGraph();
  return ITSCA_OK;
}

itsca_status itsca_get_tuning(itsca_ctx* ctx, itsca_tuning* t)
{
  if (!ctx || !t) return ITSCA_ERR_ARGUMENT;
  t->block_size = ctx->blockSize;
  t->persistent_blocks = ctx->persistentBlocks;
  t->z_subbins = ctx->zSub;
  t->phi_subrows = ctx->phiSub;
  return ITSCA_OK;
}

// Autotuner: times every tuning on the uploaded batch and keeps the fastest.
// No tuning changes a result: the chunk plans only decide which workgroup
// computes which element, and the z sub-bins and phi sub-rows only decide
// which candidates the tracklet search can skip without testing, so every
// candidate is valid.
itsca_status itsca_autotune(itsca_ctx* ctx, int repeats, itsca_tuning* chosen, float* best_ms)
{
  return guarded(ctx, "itsca_autotune", [&]() -> itsca_status {
    if (ctx->W == 0) return ctx->fail(ITSCA_ERR_STATE, "itsca_autotune: nothing uploaded");
    if (repeats < 1) repeats = 1;
    HIP_TRY(hipSetDevice(ctx->device));
    const bool graph = ctx->useGraph, prof = ctx->profiling;
    ctx->profiling = false;
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    int bestBlock = ctx->blockSize, bestBlocks = ctx->persistentBlocks, bestSub = ctx->zSub, bestPhi = ctx->phiSub;
    float best = -1.0f;
    for (int block : {64, 128, 256}) {
      for (int perCu : {4, 8, 16}) {
       for (int zSub : {1, 2, 4, 8}) {
        for (int phiSub : {1, 2, 4}) {
        if (zSub * phiSub > kMaxSubBins) continue;
        ctx->blockSize = block;
        ctx->persistentBlocks = perCu * ctx->cus;
        ctx->zSub = zSub;
        ctx->phiSub = phiSub;
        ctx->dropGraph();
        const itsca_status st = itsca_run(ctx);  // sizes the buffers, captures the graph if enabled
        if (st != ITSCA_OK) {
          (void)hipEventDestroy(a);
          (void)hipEventDestroy(b);
          return st;
        }
        HIP_TRY(hipEventRecord(a, ctx->stream));
        for (int r = 0; r < repeats; ++r) ctx->enqueue();
        HIP_TRY(hipEventRecord(b, ctx->stream));
        HIP_TRY(hipEventSynchronize(b));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, a, b));
        ms /= (float)repeats;
        if (best < 0 || ms < best) {
          best = ms;
          bestBlock = block;
          bestBlocks = ctx->persistentBlocks;
          bestSub = zSub;
          bestPhi = phiSub;
        }
        }
       }
      }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    ctx->blockSize = bestBlock;
    ctx->persistentBlocks = bestBlocks;
    ctx->zSub = bestSub;
    ctx->phiSub = bestPhi;
    ctx->profiling = prof;
    ctx->useGraph = graph;
    ctx->dropGraph();
    const itsca_status st = itsca_run(ctx);  // leave valid results of the chosen shape
    if (chosen) {
      chosen->block_size = bestBlock;
      chosen->persistent_blocks = bestBlocks;
      chosen->z_subbins = bestSub;
      chosen->phi_subrows = bestPhi;
    }
    if (best_ms) *best_ms = best;
    return st;
  });
}

// ---------------------------------------------------------------------------
// Learned tracklet filter
// ---------------------------------------------------------------------------
itsca_status itsca_mlp_init(itsca_ctx* ctx, uint64_t seed)
{
  return guarded(ctx, "itsca_mlp_init", [&]() -> itsca_status {
    HIP_TRY(hipSetDevice(ctx->device));
    // He initialisation from a splitmix64 stream (Box-Muller); zero biases.
    uint64_t st = seed;
    auto next = [&]() {
      uint64_t z = (st += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      return z ^ (z >> 31);
    };
    auto normal = [&]() {
      const double u1 = ((next() >> 11) + 1.0) * (1.0 / 9007199254740993.0);
      const double u2 = (next() >> 11) * (1.0 / 9007199254740992.0);
      return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    std::vector<float> p(kMlpParams, 0.0f);
    for (int i = 0; i < kMlpHidden * kMlpIn; ++i) p[kOffW1 + i] = (float)(normal() * std::sqrt(2.0 / kMlpIn));
    for (int i = 0; i < kMlpHidden * kMlpHidden; ++i) p[kOffW2 + i] = (float)(normal() * std::sqrt(2.0 / kMlpHidden));
    for (int i = 0; i < kMlpHidden; ++i) p[kOffW3 + i] = (float)(normal() * std::sqrt(1.0 / kMlpHidden));
    ctx->setParams(p.data());
    return ITSCA_OK;
  });
}

itsca_status itsca_mlp_set_params(itsca_ctx* ctx, const float* params)
{
  if (!ctx || !params) return ITSCA_ERR_ARGUMENT;
  return guarded(ctx, "itsca_mlp_set_params", [&]() -> itsca_status {
    HIP_TRY(hipSetDevice(ctx->device));
    ctx->setParams(params);
    return ITSCA_OK;
  });
}

itsca_status itsca_mlp_get_params(itsca_ctx* ctx, float* params)
{
  if (!ctx || !params) return ITSCA_ERR_ARGUMENT;
  return guarded(ctx, "itsca_mlp_get_params", [&]() -> itsca_status {
    HIP_TRY(hipSetDevice(ctx->device));
    ctx->ensureMlp();
    HIP_TRY(hipMemcpy(params, ctx->mlpMaster.p, kMlpParams * 4, hipMemcpyDeviceToHost));
    return ITSCA_OK;
  });
}

itsca_status itsca_set_filter(itsca_ctx* ctx, int enable, float threshold)
{
  if (!ctx) return ITSCA_ERR_ARGUMENT;
  return guarded(ctx, "itsca_set_filter", [&]() -> itsca_status {
    if (enable < 0 || enable > 2) return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_set_filter: enable must be 0, 1 or 2");
    if (enable) ctx->ensureMlp();
    ctx->filterOn = enable != 0;
    ctx->filterFused = enable == 2;
    ctx->threshold = threshold;
    ctx->dropGraph();
    return ITSCA_OK;
  });
}

itsca_status itsca_set_filter_gates(itsca_ctx* ctx, const float* gates)
{
  if (!ctx) return ITSCA_ERR_ARGUMENT;
  float g[kTrackletLayers];
  for (int l = 0; l < kTrackletLayers; ++l) {
    g[l] = gates ? gates[l] : kPhiCut;
    if (!(g[l] > 0.0f && g[l] <= kPhiCut))
      return ctx->fail(ITSCA_ERR_ARGUMENT, "itsca_set_filter_gates: every gate must lie in (0, 0.3]");
  }
  for (int l = 0; l < kTrackletLayers; ++l) ctx->phiGate[l] = g[l];
  ctx->dropGraph();
  return ITSCA_OK;
}

itsca_status itsca_train_clear(itsca_ctx* ctx, int capacity)
{
  if (!ctx || capacity < 0) return ITSCA_ERR_ARGUMENT;
  return guarded(ctx, "itsca_train_clear", [&]() -> itsca_status {
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(ctx->trainX.reserve((size_t)std::max(capacity, 1) * kMlpInPad));
    HIP_TRY(ctx->trainY.reserve(std::max(capacity, 1)));
    ctx->trainCap = capacity;
    ctx->trainN = 0;
    return ITSCA_OK;
  });
}

itsca_status itsca_train_collect(itsca_ctx* ctx, int stride, int phase, int* added)
{
  if (!ctx || stride < 1 || phase < 0) return ITSCA_ERR_ARGUMENT;
  return guarded(ctx, "itsca_train_collect", [&]() -> itsca_status {
    if (!ctx->ran) return ctx->fail(ITSCA_ERR_STATE, "itsca_train_collect: no completed run");
    if (ctx->filterOn && ctx->threshold > -1.0e38f)
      return ctx->fail(ITSCA_ERR_STATE,
                       "itsca_train_collect: labels come from plain runs, or filtered runs that keep every candidate");
    HIP_TRY(hipSetDevice(ctx->device));
    long total = 0;
    for (const Seg& g : ctx->trkSeg) total += g.count;
    HIP_TRY(ctx->trkLabel.reserve(std::max<long>(total, 1)));
    const long rows = total > phase ? (total - phase + stride - 1) / stride : 0;
    const int n = (int)std::min<long>(rows, ctx->trainCap - ctx->trainN);
    if (n > 0)
      launchLabelAndCollect(ctx->state(), (int)total, ctx->trkLabel.p, stride, phase, ctx->trainX.p, ctx->trainY.p,
                            ctx->trainN, ctx->trainCap, ctx->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->trainN += std::max(n, 0);
    if (added) *added = std::max(n, 0);
    return ITSCA_OK;
  });
}

itsca_status itsca_train_set(itsca_ctx* ctx, const uint16_t* features, const uint8_t* labels, int n)
{
  if (!ctx || n < 0 || (n > 0 && (!features || !labels))) return ITSCA_ERR_ARGUMENT;
  return guarded(ctx, "itsca_train_set", [&]() -> itsca_status {
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(ctx->trainX.reserve((size_t)std::max(n, 1) * kMlpInPad));
    HIP_TRY(ctx->trainY.reserve(std::max(n, 1)));
    if (n) {
      HIP_TRY(hipMemcpy(ctx->trainX.p, features, (size_t)n * kMlpInPad * 2, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(ctx->trainY.p, labels, (size_t)n, hipMemcpyHostToDevice));
    }
    ctx->trainN = ctx->trainCap = n;
    return ITSCA_OK;
  });
}

int itsca_train_size(const itsca_ctx* ctx) { return ctx ? ctx->trainN : -1; }

itsca_status itsca_train_get(itsca_ctx* ctx, uint16_t* features, uint8_t* labels, int n)
{
  if (!ctx || n < 0 || n > ctx->trainN) return ITSCA_ERR_ARGUMENT;
  return guarded(ctx, "itsca_train_get", [&]() -> itsca_status {
    HIP_TRY(hipSetDevice(ctx->device));
    if (n && features) HIP_TRY(hipMemcpy(features, ctx->trainX.p, (size_t)n * kMlpInPad * 2, hipMemcpyDeviceToHost));
    if (n && labels) HIP_TRY(hipMemcpy(labels, ctx->trainY.p, (size_t)n, hipMemcpyDeviceToHost));
    return ITSCA_OK;
  });
}

itsca_status itsca_train_step(itsca_ctx* ctx, int first, int count, long mult, long offset, float pos_weight, float lr,
                              float weight_decay, float* loss, float* grad)
{
  if (!ctx || first < 0 || count < 1) return ITSCA_ERR_ARGUMENT;
  return guarded(ctx, "itsca_train_step", [&]() -> itsca_status {
    if (ctx->trainN < 1) return ctx->fail(ITSCA_ERR_STATE, "itsca_train_step: empty training set");
    HIP_TRY(hipSetDevice(ctx->device));
    ctx->ensureMlp();
    if (lr > 0.0f) ++ctx->mlpStep;
    launchTrainStep(ctx->mlpBuffers(), ctx->trainX.p, ctx->trainY.p, ctx->trainN, mult, offset, first, count,
                    pos_weight, std::max(ctx->mlpStep, 1), lr, weight_decay, ctx->mlpGrad.p, ctx->mlpLoss.p,
                    ctx->stream);
    HIP_TRY(hipGetLastError());
    if (loss) HIP_TRY(hipMemcpyAsync(loss, ctx->mlpLoss.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    if (grad) HIP_TRY(hipMemcpyAsync(grad, ctx->mlpGrad.p, kMlpParams * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (lr > 0.0f) ctx->dropGraph();
    return ITSCA_OK;
  });
}

itsca_status itsca_train_epoch(itsca_ctx* ctx, int batch, float lr, float weight_decay, float pos_weight,
                               uint64_t seed, float* mean_loss)
{
  if (!ctx || batch < 1) return ITSCA_ERR_ARGUMENT;
  return guarded(ctx, "itsca_train_epoch", [&]() -> itsca_status {
    const long n = ctx->trainN;
    if (n < 1) return ctx->fail(ITSCA_ERR_STATE, "itsca_train_epoch: empty training set");
    HIP_TRY(hipSetDevice(ctx->device));
    ctx->ensureMlp();
    // pseudo-shuffle: rows (a*k + b) mod n with a prime a that does not divide n
    static const long primes[] = {1000003, 999983, 1000033, 1000037, 1000039};
    long a = 1;
    for (long pr : primes)
      if (n % pr != 0) {
        a = pr % n == 0 ? 1 : pr;
        break;
      }
    const long b = (long)(seed % (uint64_t)n);
    double sum = 0.0;
    int steps = 0;
    std::vector<float> l(1);
    for (long k0 = 0; k0 < n; k0 += batch) {
      const int count = (int)std::min<long>(batch, n - k0);
      ++ctx->mlpStep;
      launchTrainStep(ctx->mlpBuffers(), ctx->trainX.p, ctx->trainY.p, (int)n, a, b, (int)k0, count, pos_weight,
                      ctx->mlpStep, lr, weight_decay, nullptr, ctx->mlpLoss.p, ctx->stream);
      HIP_TRY(hipMemcpyAsync(l.data(), ctx->mlpLoss.p, 4, hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(hipStreamSynchronize(ctx->stream));
      sum += l[0];
      ++steps;
    }
    HIP_TRY(hipGetLastError());
    if (mean_loss) *mean_loss = (float)(sum / std::max(steps, 1));
    ctx->dropGraph();
    return ITSCA_OK;
  });
}

itsca_status itsca_train_scores(itsca_ctx* ctx, float* logits, int n)
{
  if (!ctx || !logits || n < 0 || n > ctx->trainN) return ITSCA_ERR_ARGUMENT;
  return guarded(ctx, "itsca_train_scores", [&]() -> itsca_status {
    HIP_TRY(hipSetDevice(ctx->device));
    ctx->ensureMlp();
    HIP_TRY(ctx->trainScores.reserve(std::max(n, 1)));
    if (n) {
      launchScoreRows(ctx->mlpWbf.p, ctx->trainX.p, n, ctx->trainScores.p, ctx->stream);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(logits, ctx->trainScores.p, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ITSCA_OK;
  });
}

itsca_status itsca_set_graph(itsca_ctx* ctx, int enable)
{
  if (!ctx) return ITSCA_ERR_ARGUMENT;
  ctx->useGraph = enable != 0;
  if (!ctx->useGraph) ctx->dropGraph();
  return ITSCA_OK;
}

itsca_status itsca_set_profiling(itsca_ctx* ctx, int enable)
{
  if (!ctx) return ITSCA_ERR_ARGUMENT;
  ctx->profiling = enable != 0;
  return ITSCA_OK;
}

itsca_status itsca_stage_times(itsca_ctx* ctx, float ms[5])
{
  return guarded(ctx, "itsca_stage_times", [&]() -> itsca_status {
    if (!ctx->timerReady || ctx->timer.n == 0)
      return ctx->fail(ITSCA_ERR_STATE, "itsca_stage_times: no profiled run");
    HIP_TRY(hipEventSynchronize(ctx->timer.ev[ctx->timer.n]));
    for (int k = 0; k < kNumTimedStages; ++k) ms[k] = 0.0f;
    for (int i = 0; i < ctx->timer.n; ++i) {
      float t = 0.0f;
      HIP_TRY(hipEventElapsedTime(&t, ctx->timer.ev[i], ctx->timer.ev[i + 1]));
      ms[ctx->timer.stage[i]] += t;
    }
    return ITSCA_OK;
  });
}

int itsca_kernel_times(itsca_ctx* ctx, float* ms, int cap)
{
  if (!ctx || !ctx->timerReady || ctx->timer.n == 0) return -1;
  if (hipEventSynchronize(ctx->timer.ev[ctx->timer.n]) != hipSuccess) return -1;
  const int n = std::min(cap, ctx->timer.n);
  for (int i = 0; i < n; ++i)
    if (hipEventElapsedTime(&ms[i], ctx->timer.ev[i], ctx->timer.ev[i + 1]) != hipSuccess) return -1;
  return ctx->timer.n;
}

const char* itsca_kernel_name(const itsca_ctx* ctx, int i)
{
  if (!ctx || i < 0 || i >= ctx->timer.n) return "";
  return ctx->timer.name[i];
}

}  // extern "C"
