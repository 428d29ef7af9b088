// gb_stage.h -- copies between pageable host memory and the device through
// the library's pinned staging buffers (gb_runtime.hip).
#ifndef GB_STAGE_H
#define GB_STAGE_H

#include <stdint.h>

// Both copies are synchronous and ordered after the work queued on the
// library stream.  A fill copies bytes [off, off + len) of its source into
// buf[0, len); a drain copies buf[0, len) to bytes [off, off + len) of its
// destination.  Either runs on several host threads at once over disjoint
// ranges and returns a GOOSEBERRY_ code.
typedef int (*gb_stage_fn)(void *ctx, uint64_t off, uint64_t len, char *buf);
int gb_stage_htod(char *d_dst, uint64_t bytes, gb_stage_fn fill, void *ctx);
int gb_stage_dtoh(uint64_t bytes, const char *d_src, gb_stage_fn drain, void *ctx);

#endif
