// Host-check build: roctx ranges are no-ops.
#pragma once
inline int roctxRangePushA(const char*) { return 0; }
inline int roctxRangePop() { return 0; }
