/* graphInputChannel.h -- declared in channel.h with the other channel kinds. */
#ifndef PTK_graphInputChannel_H
#define PTK_graphInputChannel_H
#include "channel.h"
#endif
