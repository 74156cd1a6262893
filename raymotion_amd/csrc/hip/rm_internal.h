// Entry points for callers inside the library (the reference-boundary shim and
// the scene C-ABI), which marshal the records themselves and take their keys
// while writing them: the same as rm_set_scene / rm_set_scene_gpu_tree, with
// the records' per-chunk key values passed in instead of hashed again. Keys
// whose chunk count does not match s->ntri are ignored (the records are then
// hashed as by the public entry points).
#pragma once

#include "rm_hip.h"
#include "rm_keys.h"

int rm_set_scene_keyed(rm_ctx* c, const rm_scene_desc* s, const rmkeys::RecordKeys& keys);
int rm_set_scene_gpu_tree_keyed(rm_ctx* c, const rm_scene_desc* s, int rebuild_request,
                                const rmkeys::RecordKeys& keys);
