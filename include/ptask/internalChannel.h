/* internalChannel.h -- declared in channel.h with the other channel kinds. */
#ifndef PTK_internalChannel_H
#define PTK_internalChannel_H
#include "channel.h"
#endif
