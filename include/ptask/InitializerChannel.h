/* InitializerChannel.h -- declared in channel.h with the other channel kinds. */
#ifndef PTK_InitializerChannel_H
#define PTK_InitializerChannel_H
#include "channel.h"
#endif
