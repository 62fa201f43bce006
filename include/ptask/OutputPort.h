/* OutputPort.h -- declared in port.h with the other port kinds. */
#ifndef PTK_OutputPort_H
#define PTK_OutputPort_H
#include "port.h"
#endif
