// Entry points for callers inside the library (the reference-boundary shim and
// the scene C-ABI), which marshal the records themselves and take their keys
// while writing them: the same as rm_set_scene / rm_set_scene_gpu_tree, with
// the records' per-chunk key values passed in instead of hashed again. Keys
// whose chunk count does not match s->ntri are ignored (the records are then
// hashed as by the public entry points).
#pragma once

#include "rm_hip.h"
#include "rm_keys.h"
#include "rm_pack.h"

int rm_set_scene_keyed(rm_ctx* c, const rm_scene_desc* s, const rmkeys::RecordKeys& keys);
int rm_set_scene_gpu_tree_keyed(rm_ctx* c, const rm_scene_desc* s, int rebuild_request,
                                const rmkeys::RecordKeys& keys, bool staged = false);

// Geometry staging, for callers that marshal scenes whose vertices change
// from call to call. rm_geo_stage_begin sizes device staging buffers for ntri
// triangles; rm_geo_stage_put queues the copy of the device-format geometry of
// triangles [first, first + count) (scene order, as rmh::pack_tri_geo and
// pack_tri_nrm write it, in pinned memory left unchanged until the upload
// returns) into them, and may run on several threads at once while the
// caller marshals. rm_set_scene_gpu_tree_keyed with staged = true then takes
// the staged geometry in place of packing and copying it, if all ntri
// triangles were staged without error (else it packs and copies as usual).
int rm_geo_stage_begin(rm_ctx* c, size_t ntri);
bool rm_geo_stage_put(rm_ctx* c, size_t first, size_t count, const rmh::F4* geo, const rmh::F4* nrm,
                      const rmh::F4* nn);
