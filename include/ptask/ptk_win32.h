/*
 * ptk_win32.h -- the Win32 vocabulary types that the PTask public API is written in, as plain C
 * types for Linux, so that application code written against the reference headers compiles
 * unchanged.  Nothing here wraps an operating-system service.
 */
#ifndef PTK_WIN32_H
#define PTK_WIN32_H

#include <stddef.h>
#include <stdint.h>

typedef int BOOL;
typedef unsigned int UINT;
typedef unsigned long DWORD;
typedef long LONG;
typedef unsigned long ULONG;
typedef long long LONGLONG;
typedef unsigned long long ULONGLONG;
typedef void VOID;
typedef void *HANDLE;
typedef void *LPVOID;
typedef char CHAR;
typedef wchar_t WCHAR;
typedef long HRESULT;
typedef void (*FARPROC)(void);

#ifndef TRUE
#define TRUE 1
#endif
#ifndef FALSE
#define FALSE 0
#endif
#ifndef INFINITE
#define INFINITE 0xFFFFFFFFul
#endif
#ifndef S_OK
#define S_OK ((HRESULT)0L)
#endif
#ifndef E_FAIL
#define E_FAIL ((HRESULT)0x80004005L)
#endif
#ifndef SUCCEEDED
#define SUCCEEDED(hr) (((HRESULT)(hr)) >= 0)
#endif
#ifndef FAILED
#define FAILED(hr) (((HRESULT)(hr)) < 0)
#endif

/* Calling-convention keywords carry no meaning on this platform.  (SAL parameter annotations
 * such as __in are not defined: those names are used inside the C++ standard library.) */
#ifndef __stdcall
#define __stdcall
#endif
#ifndef __declspec
#define __declspec(x)
#endif
#ifndef __int64
#define __int64 long long
#endif

#endif /* PTK_WIN32_H */
