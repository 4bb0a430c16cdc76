/* fvgs_server: many independent SLAM sessions served from one GPU.
 *
 * No counterpart in the reference, which runs one GSSlam instance per process. The server keeps
 * each session's map resident in device memory, admits sessions against a device-memory budget,
 * and runs their tracking and mapping requests on `lanes` concurrent request slots (one HIP
 * stream and one worker thread each; every lane owns a fvgs_ctx). A request submitted while
 * others are on the GPU starts as soon as a lane is free: requests of different sessions overlap
 * on the device, and each request's result is the same as when it runs alone (tracking results
 * bit for bit; mapping gradients up to the order of the float atomics of the backward).
 *
 * All functions return FVGS_OK (0) or a negative FVGS_E_* code (fvgs.h).
 */
#ifndef FVGS_SERVER_H
#define FVGS_SERVER_H

#include <stddef.h>
#include <stdint.h>

#include "fvgs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fvgs_server fvgs_server;

typedef struct fvgs_server_params {
    int lanes;              /* concurrent request slots (4) */
    int max_sessions;       /* (64) */
    size_t hbm_budget;      /* device bytes for maps, frames, keyframes and lane workspaces
                               (0 = 200 GiB) */
    fvgs_params params;     /* per-lane constants (fvgs_default_params) */
} fvgs_server_params;

typedef enum fvgs_request_kind { FVGS_REQ_TRACK = 1, FVGS_REQ_MAPPING = 2 } fvgs_request_kind;

typedef struct fvgs_request_result {
    int status;             /* FVGS_OK or the error of the request */
    int kind;               /* fvgs_request_kind */
    int session;
    int iters;              /* tracking: iterations run */
    double pose7[7];        /* tracking: final pose (px, py, pz, qx, qy, qz, qw) */
    double queue_ms;        /* submit -> start on a lane */
    double run_ms;          /* start -> end (host wall clock) */
    double device_ms;       /* tracking: sum of the per-iteration device times */
} fvgs_request_result;

int fvgs_server_default_params(fvgs_server_params* p);
int fvgs_server_create(const fvgs_server_params* p, fvgs_server** out);
/* Waits for queued and running requests, then frees everything. */
int fvgs_server_destroy(fvgs_server* srv);
const char* fvgs_server_last_error(const fvgs_server* srv);

/* Admits a session: copies its map (host or device arrays, as fvgs_set_gaussians) into server
 * memory together with a zeroed gradient accumulator. frame_w x frame_h is the largest frame the
 * session will submit; the lanes' workspaces are sized for the largest admitted session.
 * Returns FVGS_E_OOM ("rejected") and allocates nothing when the session would not fit in the
 * budget, FVGS_E_STATE when max_sessions are open. */
int fvgs_server_open_session(fvgs_server* srv, uint32_t n, const float* pos3, const float* scale3,
                             const float* quat4, const float* color3, const float* opacity,
                             int mem, int frame_w, int frame_h, int* session);
/* Waits for the session's requests, then frees its memory. */
int fvgs_server_close_session(fvgs_server* srv, int session);

/* Stores a keyframe (pose, colour and depth images) on the device for mapping requests. */
int fvgs_server_add_keyframe(fvgs_server* srv, int session, const fvgs_pose* pose, int width,
                             int height, const float* rgb3, const float* depth, int mem,
                             int* keyframe);

/* Tracking request: iters device-resident Gauss-Newton iterations (fvgs_track) from pose7
 * against the frame (rgb3, depth; gradients are computed on the device). Host frame buffers are
 * read when the request starts and must stay valid until it has completed. */
int fvgs_server_submit_track(fvgs_server* srv, int session, int width, int height,
                             const float* rgb3, const float* depth, int mem,
                             const fvgs_intrinsics* intr, int iters, const double pose7[7],
                             uint64_t* ticket);
/* Mapping request: fvgs_mapping_step on each listed keyframe; the views' 3D gradients are added
 * into the session's accumulator. A session's mapping requests run one at a time. */
int fvgs_server_submit_mapping(fvgs_server* srv, int session, const fvgs_intrinsics* intr,
                               int n_keyframes, const int* keyframes, uint64_t* ticket);
/* Blocks until the request has completed; fills *out and forgets the ticket. */
int fvgs_server_wait(fvgs_server* srv, uint64_t ticket, fvgs_request_result* out);

/* Copies the session's accumulated 3D gradients (fvgs_delta3d[n]) to host memory. */
int fvgs_server_read_gradients(fvgs_server* srv, int session, fvgs_delta3d* host, size_t bytes);
/* Device bytes held by sessions and by lane workspaces (reserved), and the budget. */
int fvgs_server_memory(fvgs_server* srv, size_t* sessions, size_t* lanes, size_t* budget);

#ifdef __cplusplus
}
#endif

#endif /* FVGS_SERVER_H */
