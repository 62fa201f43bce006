/* multichannel.h -- declared in channel.h with the other channel kinds. */
#ifndef PTK_multichannel_H
#define PTK_multichannel_H
#include "channel.h"
#endif
