// Host-side raw sample file I/O and the context-free error channel.
#ifndef MEMD_BINIO_H
#define MEMD_BINIO_H

namespace memd {
// Message of the last failure of a call made without a context (thread-local).
const char* global_error();
void set_global_error(const char* msg);
}  // namespace memd

#endif
