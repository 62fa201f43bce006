/* MetaPort.h -- declared in port.h with the other port kinds. */
#ifndef PTK_MetaPort_H
#define PTK_MetaPort_H
#include "port.h"
#endif
