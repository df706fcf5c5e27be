// apps/apps.h -- the apps compiled into the library.  Each entry names an
// app's boundary policy (apps/<app>/boundary.h); the step kernel builds its
// march once per entry and runs, per work item, the one of its instance's
// boundary kind.  A new app: its directory, its include below, and one
// FIVONAGI_APP line.
#ifndef FIVONAGI_APPS_APPS_H
#define FIVONAGI_APPS_APPS_H

#include "common/boundary.h"
#include "edge-drive/boundary.h"
#include "hifu-beam/boundary.h"
#include "taylor-angle-xy/boundary.h"

// X(policy): every registered boundary policy
#define FIVONAGI_APPS(X) \
  X(fvbc::ZeroGradient)  \
  X(fvbc::EdgeDrive)     \
  X(fvbc::Taylor)        \
  X(fvbc::Hifu)

#endif  // FIVONAGI_APPS_APPS_H
