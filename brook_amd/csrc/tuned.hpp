// Launch shapes and frontier chunk sizes of the persistent label kernels. brook_amd/tune.py rewrites
// this file with the values it measures fastest on the device it runs on (a -D on the compiler
// line still overrides each one). Every value gives bit-identical results; only speed changes.
#pragma once
// Last tuned: {"batch_ms": 112.91, "device": "gfx950", "plain_ms": 184.82}
#ifndef BROOK_TRACE_BLOCK
#define BROOK_TRACE_BLOCK 1024  // threads per trace_labels workgroup (one label at a time)
#endif
#ifndef BROOK_NB_BATCH
#define BROOK_NB_BATCH 9  // neighbours a trace thread keeps in flight together (9, 13 or 26)
#endif
#ifndef BROOK_FLOOD_BLOCK
#define BROOK_FLOOD_BLOCK 1024  // threads per prepare_labels workgroup
#endif
#ifndef BROOK_FLOOD_NB_BATCH
#define BROOK_FLOOD_NB_BATCH 9  // neighbours a flood thread keeps in flight together (9, 13 or 26)
#endif
#ifndef BROOK_TRACE_GRID
#define BROOK_TRACE_GRID 0  // trace_labels workgroups resident at once (0: two per compute unit)
#endif
#ifndef BROOK_FLOOD_GRID
#define BROOK_FLOOD_GRID 0  // prepare_labels workgroups resident at once (0: two per compute unit)
#endif
