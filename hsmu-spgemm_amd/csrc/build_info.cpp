// build_info.cpp -- which sources this libhsmu.so was built from.  The Makefile passes the
// last git commit that changed csrc/ or include/ (with "+dirty" when they differ from it)
// and a hash of every source and header file, so a test or benchmark run can show that
// the library it loaded was built from the sources beside it (hsmu.source_hash()
// recomputes the same hash from the tree).
#include "../../include/hsmu.h"

#ifndef HSMU_BUILD_HEAD
#define HSMU_BUILD_HEAD "unknown"
#endif
#ifndef HSMU_SRC_HASH
#define HSMU_SRC_HASH "unknown"
#endif

extern "C" const char *hsmu_build_info(void) { return "src_commit=" HSMU_BUILD_HEAD " src=" HSMU_SRC_HASH; }
