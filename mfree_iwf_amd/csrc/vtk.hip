// Legacy VTK output of the bodies' state, the reference's results files
// (src/vtk_writer.h:27-127, written every 10th step by src/mfree_iwf.cpp:
// 82,99-103): an ASCII unstructured grid of one vertex cell per particle with
// the point data h, density, Sxx, Sxy, Syy and the vector velocity, bodies in
// order, particles in idx order, every value printed with %f.
//
// The context writer snapshots the state on the device (one gather per field
// into idx order, into a staging buffer of its own), copies the snapshot to
// pinned host memory on a copy stream, and formats the file on a host thread:
// the context's stream only runs the gathers, so stepping goes on while the
// copy and the formatting run.  Formatting is split over host threads by
// particle range; the pieces are written in order.
#include "../../include/mfree_iwf_amd.h"

#include <stdio.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "ctx.h"

namespace mfree {

// Snapshot fields, in the order the file uses them.
enum : int { VX_X, VX_Y, VX_H, VX_RHO, VX_SXX, VX_SXY, VX_SYY, VX_VX, VX_VY, kVtkFields };

struct VtkView {
  const double* f[kVtkFields];
  uint64_t n;
};

struct VtkState {
  double* stage = nullptr;  // device, kVtkFields * N (in the context's allocations)
  double* host = nullptr;   // pinned, kVtkFields * N
  hipStream_t copy = nullptr;
  hipEvent_t gathered = nullptr, copied = nullptr;
  std::thread worker;
  bool pending = false;
  mfree_status result = MFREE_OK;
  std::string msg;
};

namespace {

// Particles formatted per piece of work (one host thread formats a range of
// every section).
constexpr uint64_t kVtkChunk = 1u << 16;

void put(std::string& s, const char* fmt, double a) {
  char b[64];
  const int k = snprintf(b, sizeof b, fmt, a);
  s.append(b, (size_t)k);
}

void put3(std::string& s, double a, double b) {
  char t[160];
  const int k = snprintf(t, sizeof t, "%f %f %f\n", a, b, 0.);
  s.append(t, (size_t)k);
}

// Per-particle sections of the file: points, cells, cell types, h, density,
// Sxx, Sxy, Syy, velocity.
constexpr int kSections = 9;

void format_range(const VtkView& v, uint64_t b, uint64_t e, std::string* out /* kSections */) {
  for (uint64_t i = b; i < e; i++) put3(out[0], v.f[VX_X][i], v.f[VX_Y][i]);
  for (uint64_t i = b; i < e; i++) {
    char t[32];
    const int k = snprintf(t, sizeof t, "%d %d\n", 1, (int)i);
    out[1].append(t, (size_t)k);
  }
  out[2].reserve(2 * (e - b));
  for (uint64_t i = b; i < e; i++) out[2].append("1\n", 2);
  const int scal[5] = {VX_H, VX_RHO, VX_SXX, VX_SXY, VX_SYY};
  for (int s = 0; s < 5; s++)
    for (uint64_t i = b; i < e; i++) put(out[3 + s], "%f\n", v.f[scal[s]][i]);
  for (uint64_t i = b; i < e; i++) put3(out[8], v.f[VX_VX][i], v.f[VX_VY][i]);
}

mfree_status write_file(const char* path, const VtkView& v, std::string& msg) {
  const uint64_t n = v.n;
  const uint64_t nchunks = std::max<uint64_t>(1, (n + kVtkChunk - 1) / kVtkChunk);
  std::vector<std::string> piece(nchunks * kSections);
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const uint64_t nthreads = std::min<uint64_t>({nchunks, (uint64_t)hw, 16});
  auto work = [&](uint64_t t) {
    for (uint64_t c = t; c < nchunks; c += nthreads)
      format_range(v, c * kVtkChunk, std::min(n, (c + 1) * kVtkChunk), &piece[c * kSections]);
  };
  std::vector<std::thread> pool;
  for (uint64_t t = 1; t < nthreads; t++) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();

  FILE* fp = fopen(path, "w+");
  if (!fp) {
    msg = std::string("cannot open ") + path;
    return MFREE_ERR_ARG;
  }
  const int ni = (int)n;
  // section headers: the lines before each per-particle section
  char head[kSections][160];
  snprintf(head[0], sizeof head[0],
           "# vtk DataFile Version 2.0\nmfree iwf\nASCII\n\nDATASET UNSTRUCTURED_GRID\nPOINTS %d float\n", ni);
  snprintf(head[1], sizeof head[1], "\nCELLS %d %d\n", ni, 2 * ni);
  snprintf(head[2], sizeof head[2], "\nCELL_TYPES %d\n", ni);
  snprintf(head[3], sizeof head[3], "\nPOINT_DATA %d\nSCALARS h float 1\nLOOKUP_TABLE default\n", ni);
  snprintf(head[4], sizeof head[4], "\nSCALARS density float 1\nLOOKUP_TABLE default\n");
  snprintf(head[5], sizeof head[5], "\nSCALARS Sxx float 1\nLOOKUP_TABLE default\n");
  snprintf(head[6], sizeof head[6], "\nSCALARS Sxy float 1\nLOOKUP_TABLE default\n");
  snprintf(head[7], sizeof head[7], "\nSCALARS Syy float 1\nLOOKUP_TABLE default\n");
  snprintf(head[8], sizeof head[8], "\nVECTORS velocity float\n");
  bool ok = true;
  for (int s = 0; s < kSections && ok; s++) {
    ok = fputs(head[s], fp) >= 0;
    for (uint64_t c = 0; c < nchunks && ok; c++) {
      const std::string& p = piece[c * kSections + s];
      ok = fwrite(p.data(), 1, p.size(), fp) == p.size();
    }
  }
  ok = ok && fputs("\n", fp) >= 0;
  ok = (fclose(fp) == 0) && ok;
  if (!ok) {
    msg = std::string("write failed: ") + path;
    return MFREE_ERR_ARG;
  }
  return MFREE_OK;
}

mfree_status vtk_finish(mfree_ctx* c) {
  VtkState* v = c->vtk;
  if (!v || !v->pending) return MFREE_OK;
  v->worker.join();
  v->pending = false;
  if (v->result != MFREE_OK) return fail(c, v->result, "VTK: " + v->msg);
  return MFREE_OK;
}

}  // namespace

void vtk_free(mfree_ctx* c) {
  VtkState* v = c->vtk;
  if (!v) return;
  if (v->pending) v->worker.join();
  if (v->host) (void)hipHostFree(v->host);
  if (v->gathered) (void)hipEventDestroy(v->gathered);
  if (v->copied) (void)hipEventDestroy(v->copied);
  if (v->copy) (void)hipStreamDestroy(v->copy);
  delete v;
  c->vtk = nullptr;
}

}  // namespace mfree

using namespace mfree;

extern "C" {

mfree_status mfree_vtk_write_host(const char* path, uint32_t nbodies, const mfree_particles* bodies,
                                  const uint32_t* counts) {
  if (!path || (nbodies && (!bodies || !counts))) return MFREE_ERR_ARG;
  uint64_t n = 0;
  for (uint32_t k = 0; k < nbodies; k++) {
    const mfree_particles& p = bodies[k];
    if (counts[k] && (!p.x || !p.y || !p.h || !p.rho || !p.Sxx || !p.Sxy || !p.Syy || !p.vx || !p.vy))
      return MFREE_ERR_ARG;
    n += counts[k];
  }
  if (n >= (1ull << 30)) return MFREE_ERR_ARG;
  // bodies concatenated (one copy per field; a single body is used in place)
  std::vector<double> cat;
  VtkView v;
  v.n = n;
  if (nbodies == 1) {
    const mfree_particles& p = bodies[0];
    const double* f[kVtkFields] = {p.x, p.y, p.h, p.rho, p.Sxx, p.Sxy, p.Syy, p.vx, p.vy};
    for (int k = 0; k < kVtkFields; k++) v.f[k] = f[k];
  } else {
    cat.resize(kVtkFields * n);
    uint64_t o = 0;
    for (uint32_t b = 0; b < nbodies; b++) {
      const mfree_particles& p = bodies[b];
      const double* f[kVtkFields] = {p.x, p.y, p.h, p.rho, p.Sxx, p.Sxy, p.Syy, p.vx, p.vy};
      for (int k = 0; k < kVtkFields; k++)
        if (counts[b]) std::copy(f[k], f[k] + counts[b], cat.begin() + k * n + o);
      o += counts[b];
    }
    for (int k = 0; k < kVtkFields; k++) v.f[k] = cat.data() + k * n;
  }
  std::string msg;
  return write_file(path, v, msg);
}

mfree_status mfree_vtk_write_async(mfree_ctx* c, const char* path) {
  if (!c || !path) return MFREE_ERR_ARG;
  mfree_status st = commit(c);
  if (st) return st;
  HIPCHK(c, hipSetDevice(c->opt.device));
  st = vtk_finish(c);  // one file in flight: its buffers are reused below
  if (st) return st;
  const size_t N = c->N;
  if (!c->vtk) {
    VtkState* v = new VtkState();
    c->vtk = v;
    HIPCHK(c, dalloc(c, &v->stage, kVtkFields * N));
    HIPCHK(c, hipHostMalloc((void**)&v->host, kVtkFields * N * sizeof(double), hipHostMallocDefault));
    HIPCHK(c, hipStreamCreateWithFlags(&v->copy, hipStreamNonBlocking));
    HIPCHK(c, hipEventCreateWithFlags(&v->gathered, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&v->copied, hipEventDisableTiming));
  }
  VtkState* v = c->vtk;
  const int b = c->cur;
  const double* src[kVtkFields] = {c->y[b].f[0], c->y[b].f[1], c->h[b],      c->y[b].f[4], c->y[b].f[5],
                                    c->y[b].f[6], c->y[b].f[7], c->y[b].f[2], c->y[b].f[3]};
  for (int k = 0; k < kVtkFields; k++)
    launch_to_idx_order(c->stream, (uint32_t)N, c->gid[b], src[k], v->stage + k * N);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(v->gathered, c->stream));
  HIPCHK(c, hipStreamWaitEvent(v->copy, v->gathered, 0));
  HIPCHK(c, hipMemcpyAsync(v->host, v->stage, kVtkFields * N * sizeof(double), hipMemcpyDeviceToHost, v->copy));
  HIPCHK(c, hipEventRecord(v->copied, v->copy));
  v->pending = true;
  v->result = MFREE_OK;
  v->msg.clear();
  const std::string file(path);
  const int dev = c->opt.device;
  v->worker = std::thread([v, file, N, dev]() {
    (void)hipSetDevice(dev);
    if (hipEventSynchronize(v->copied) != hipSuccess) {
      v->result = MFREE_ERR_HIP;
      v->msg = "snapshot copy failed";
      return;
    }
    VtkView view;
    view.n = N;
    for (int k = 0; k < kVtkFields; k++) view.f[k] = v->host + k * N;
    v->result = write_file(file.c_str(), view, v->msg);
  });
  return MFREE_OK;
}

mfree_status mfree_vtk_wait(mfree_ctx* c) {
  if (!c) return MFREE_ERR_ARG;
  return vtk_finish(c);
}

mfree_status mfree_vtk_write(mfree_ctx* c, const char* path) {
  mfree_status st = mfree_vtk_write_async(c, path);
  if (st) return st;
  return vtk_finish(c);
}

}  // extern "C"
