/* InputPort.h -- declared in port.h with the other port kinds. */
#ifndef PTK_InputPort_H
#define PTK_InputPort_H
#include "port.h"
#endif
