/* graphOutputChannel.h -- declared in channel.h with the other channel kinds. */
#ifndef PTK_graphOutputChannel_H
#define PTK_graphOutputChannel_H
#include "channel.h"
#endif
