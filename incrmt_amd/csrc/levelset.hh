// Narrow-band level set of one solid on the GPU: pinning to the reference
// map, band rebuild (reinitialisation) and reference-map extension.
#ifndef INCRMT_LEVELSET_HH
#define INCRMT_LEVELSET_HH

#include <vector>
#include "common.hh"
#include "eikonal.hh"

namespace incrmt {

// Counter slots in LsWork::cnt
enum { kCntIface = 0, kCntCand, kCntPband, kCntErr, kCntTileFim, kCntTileExt, kCntIter };
constexpr int kIterSlots = 64;
enum { kErrNewton = 1, kErrSingular = 2 };

// The fixed-point passes run on kTileX x kTileY tiles of the ghosted grid
// held in LDS, one workgroup of kTileX * kTileY threads per tile.
#ifndef INCRMT_LS_TILE_X
#define INCRMT_LS_TILE_X 16
#endif
#ifndef INCRMT_LS_TILE_Y
#define INCRMT_LS_TILE_Y 16
#endif
constexpr int kTileX = INCRMT_LS_TILE_X, kTileY = INCRMT_LS_TILE_Y, kTileThreads = kTileX * kTileY;
static_assert(kTileThreads % 64 == 0 && kTileThreads <= 1024, "a tile is whole wavefronts of one workgroup");

struct LsWork {
    double *va, *vb;      // fixed-point buffers (ghosted grid)
    double *ex, *ey;      // extension buffers (ghosted grid), ping
    double *fx, *fy;      // extension buffers, pong
    int *iface, *cand, *pband;  // point lists
    unsigned char *mark;  // dilation scratch
    unsigned char *cflag; // 1 where a point takes part in the march
    int *tflag;           // per-tile flags: bit 1 march, bit 2 extension
    int *tiles_fim, *tiles_ext;  // active tile lists
    int ntx, nty;         // tile grid
    int *cnt;             // counters (kCntIter + kIterSlots ints)
    int *hcnt;            // pinned host mirror
    int pred_fim = 2, pred_ext = 2;  // passes the last march / extension needed
    // window of the next rebuild (see rebuild_common): level-set points
    // [win0,win1) x [win2,win3); bbox of the last band on the device, its
    // host copy and the event that marks the copy done
    bool have_win = false, bbox_pending = false, pin_first = false;
    int win[4] = {0, 0, 0, 0};
    int *bbox = nullptr, *hbbox = nullptr;
    hipEvent_t ev_bbox = nullptr;
};

int ls_alloc(const Grid &g, LsWork &w, hipStream_t s);
void ls_free(LsWork &w);

// Rebuilds and extensions of several solids of one grid run their passes
// together: one launch per pass for every solid, one counter read-back per
// batch of passes. Every solid gets exactly what the single-solid call
// gives it. The batch work area holds up to `cap` solids.
struct LsJob {
    const ObjConst *c;
    const ObjArrays *a;
    LsWork *w;
};
struct LsBatch {
    void *d_items = nullptr, *h_items = nullptr;
    int *d_cnt = nullptr, *h_cnt = nullptr;
    int cap = 0;
    // per-solid items of the shared set-up, finish and copy launches: four
    // regions (set-up, finish, copy-in, copy-out) in pinned memory and on
    // the device, each rewritten only after its last upload has run
    void *d_setup = nullptr, *h_setup = nullptr;
    size_t setup_stride = 0;
    hipEvent_t ev_setup[4] = {nullptr, nullptr, nullptr, nullptr};
    bool setup_used[4] = {false, false, false, false};
};
int ls_batch_alloc(int cap, LsBatch &bt);
void ls_batch_free(LsBatch &bt);
// ls_pin_rebuild of every job; *fim_iters: the most passes any took
int ls_pin_rebuild_batch(hipStream_t s, const Grid &g, const std::vector<LsJob> &jobs, LsBatch &bt,
                         int *fim_iters);
// ls_pin_rebuild_batch in two halves (as ls_pin_rebuild_begin/_end):
// begin enqueues the set-up and the predicted batch of march passes on s
// and returns the pending state; end waits for s, runs any further passes
// and the finish, and frees the state. pin_new: pin to (nX, nY) on the
// interior cells instead of (X, Y). After an error between the halves,
// synchronise s and free the state with ls_batch_pending_free.
struct LsBatchPending;
int ls_pin_rebuild_batch_begin(hipStream_t s, const Grid &g, const std::vector<LsJob> &jobs, LsBatch &bt,
                               bool pin_new, LsBatchPending **out);
int ls_pin_rebuild_batch_end(hipStream_t s, const Grid &g, const std::vector<LsJob> &jobs, LsBatch &bt,
                             LsBatchPending *p, int *fim_iters);
void ls_batch_pending_free(LsBatchPending *p);
// ls_extend of every job; *ext_iters: the most passes any took
int ls_extend_batch(hipStream_t s, const Grid &g, const std::vector<LsJob> &jobs, LsBatch &bt, int *ext_iters);

// obj_field::pin + levelset::build_band (obj_field.cc:92-97,
// levelset.cc:199-335). Returns 0, or an INCRMT_ERR_* code.
int ls_pin_rebuild(hipStream_t s, const Grid &g, const ObjConst &c, const ObjArrays &a, LsWork &w,
                   int *fim_iters);
// ls_pin_rebuild in two halves, for a caller that runs it beside other
// work on another stream s. begin enqueues the set-up and the predicted
// batch of march passes and returns without waiting; end waits for s, runs
// any further passes and the finish, as ls_pin_rebuild does. (pin_x,
// pin_y), when given, stand in for (X, Y) on the interior cells in the pin.
// Rectangle of level-set points [I0,I1) x [J0,J1) a rebuild processes.
struct Win {
    int I0, I1, J0, J1;
};
struct LsRebuild {
    Win wn{};
    bool whole = false, live = false;
    int passes = 0, batch = 0;
    EikGeom eg{};
};
int ls_pin_rebuild_begin(hipStream_t s, const Grid &g, const ObjConst &c, const ObjArrays &a, LsWork &w,
                         const double *pin_x, const double *pin_y, LsRebuild &r);
int ls_pin_rebuild_end(hipStream_t s, const Grid &g, const ObjConst &c, const ObjArrays &a, LsWork &w, LsRebuild &r,
                       int *fim_iters);
// Make the next rebuild cover the whole grid (after phi or the status
// codes were replaced from outside).
void ls_reset_window(LsWork &w);
// levelset::build_band alone on the current phi (initialisation).
int ls_rebuild(hipStream_t s, const Grid &g, const ObjConst &c, const ObjArrays &a, LsWork &w, int *fim_iters);
// obj_field::extrapolate (traverse.hh:169-260) over the band of the last
// rebuild. Returns 0, or an INCRMT_ERR_* code.
int ls_extend(hipStream_t s, const Grid &g, const ObjArrays &a, LsWork &w, int *ext_iters);

}  // namespace incrmt

#endif
