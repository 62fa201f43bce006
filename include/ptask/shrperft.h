/* shrperft.h -- the shared timer shares CHighResolutionTimer. */
#ifndef PTK_SHRPERFT_H
#define PTK_SHRPERFT_H
#include "hrperft.h"
#endif
