// Launch shapes and frontier chunk sizes of the persistent label kernels. brook_amd/tune.py rewrites
// this file with the values it measures fastest on the device it runs on (a -D on the compiler
// line still overrides each one). Every value gives bit-identical results; only speed changes.
#pragma once
// Last tuned: {"batch_ms": 111.79, "device": "MI355X", "plain_ms": 192.83}
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
