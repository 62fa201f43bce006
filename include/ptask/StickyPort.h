/* StickyPort.h -- declared in port.h with the other port kinds. */
#ifndef PTK_StickyPort_H
#define PTK_StickyPort_H
#include "port.h"
#endif
