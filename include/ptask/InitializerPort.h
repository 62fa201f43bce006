/* InitializerPort.h -- declared in port.h with the other port kinds. */
#ifndef PTK_InitializerPort_H
#define PTK_InitializerPort_H
#include "port.h"
#endif
