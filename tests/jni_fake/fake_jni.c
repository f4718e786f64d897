/*
 * Fake JNIEnv / JavaVM for testing the JNI binding of
 * libNativeSPIMReconstructionCuda.so without a JVM (this image has no JDK).
 *
 * Implements, at the slot numbers of the JNI specification's function
 * tables, the calls the binding makes: FindClass, ThrowNew, ExceptionCheck,
 * New/DeleteGlobalRef, DeleteLocalRef, GetStaticMethodID, CallStaticObjectMethod,
 * CallStaticVoidMethod, NewStringUTF, Get/ReleaseStringUTFChars, GetArrayLength,
 * GetObjectArrayElement, New{Byte,Short}Array, Get/Release{Byte,Short,Float}
 * ArrayElements, Set{Byte,Short}ArrayRegion, GetJavaVM,
 * Get/ReleasePrimitiveArrayCritical and JavaVM::AttachCurrentThread.  Every
 * other slot points at a stub that counts unexpected calls.
 *
 * The "calling class" has two static methods: getNextPlane() returns the
 * next plane (an Object[] of nViews short[]/byte[]) from planes installed
 * with fj_set_planes, then null; returnNextPlane(Object) records a copy.
 */
#include <stdarg.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { K_CLASS = 1, K_STRING, K_OBJARR, K_SHORTARR, K_BYTEARR, K_FLOATARR };

typedef struct Obj {
    int kind;
    int len;
    void *data;      /* payload (string chars, primitive data, Obj* array) */
    struct Obj *next; /* allocation list */
} Obj;

static void *g_table[240];
static void *const *g_env = NULL;  /* JNIEnv value: pointer to the table */
static void *g_vm_table[8];
static void *const *g_vm = NULL;

static Obj *g_objs = NULL;
static Obj g_class = {K_CLASS, 0, NULL, NULL};
static Obj g_exc_class = {K_CLASS, 0, NULL, NULL};
static char g_exception[512];
static int g_pending = 0;
static int g_unexpected = 0;
static int g_local_deleted = 0;
static int g_global_refs = 0;
static int g_methods_present = 1;

static int g_kind = 16, g_nplanes = 0, g_nviews = 0, g_planesize = 0, g_next = 0;
static unsigned char *g_planes = NULL;
static int g_nret = 0, g_ret_cap = 0;
static unsigned char **g_ret = NULL;

static int MID_GET = 1, MID_RET = 2;

/* upcall log: 'G' a served plane, 'g' end of data (null), 'R' a returned plane */
static char g_log[8192];
static int g_loglen = 0;
static void log_call(char c)
{
    if (g_loglen < (int)sizeof g_log - 1) g_log[g_loglen++] = c;
    g_log[g_loglen] = 0;
}
/* a served plane whose view `g_short_view` has only `g_short_len` samples */
static int g_short_plane = -1, g_short_view = 0, g_short_len = 0;

static Obj *new_obj(int kind, int len, size_t bytes)
{
    Obj *o = (Obj *)calloc(1, sizeof(Obj));
    o->kind = kind;
    o->len = len;
    o->data = calloc(bytes ? bytes : 1, 1);
    o->next = g_objs;
    g_objs = o;
    return o;
}

static size_t elem_size(int kind)
{
    switch (kind) {
    case K_SHORTARR: return 2;
    case K_BYTEARR: return 1;
    case K_FLOATARR: return 4;
    case K_OBJARR: return sizeof(Obj *);
    default: return 1;
    }
}

/* ---- JNIEnv functions ---- */
static void unexpected(void) { g_unexpected++; }
static void *FindClass(void *env, const char *name) { (void)env; (void)name; return &g_exc_class; }
static int ThrowNew(void *env, void *cls, const char *msg)
{
    (void)env; (void)cls;
    strncpy(g_exception, msg ? msg : "", sizeof g_exception - 1);
    g_pending = 1;
    return 0;
}
static uint8_t ExceptionCheck(void *env) { (void)env; return (uint8_t)g_pending; }
static void *NewGlobalRef(void *env, void *o) { (void)env; g_global_refs++; return o; }
static void DeleteGlobalRef(void *env, void *o) { (void)env; (void)o; g_global_refs--; }
static void DeleteLocalRef(void *env, void *o) { (void)env; (void)o; g_local_deleted++; }
static void *GetStaticMethodID(void *env, void *cls, const char *name, const char *sig)
{
    (void)env; (void)cls;
    if (!g_methods_present) {
        strncpy(g_exception, "java/lang/NoSuchMethodError", sizeof g_exception - 1);
        g_pending = 1;
        return NULL;
    }
    if (!strcmp(name, "getNextPlane") && !strcmp(sig, "()[Ljava/lang/Object;")) return &MID_GET;
    if (!strcmp(name, "returnNextPlane") && !strcmp(sig, "(Ljava/lang/Object;)V")) return &MID_RET;
    return NULL;
}
static void *CallStaticObjectMethod(void *env, void *cls, void *mid, ...)
{
    (void)env; (void)cls;
    if (mid != &MID_GET) return NULL;
    if (g_next >= g_nplanes) {
        log_call('g');
        return NULL;
    }
    log_call('G');
    const int ak = g_kind == 16 ? K_SHORTARR : K_BYTEARR;
    const size_t es = g_kind == 16 ? 2 : 1;
    Obj *arr = new_obj(K_OBJARR, g_nviews, sizeof(Obj *) * g_nviews);
    for (int v = 0; v < g_nviews; ++v) {
        const int len = (g_next == g_short_plane && v == g_short_view) ? g_short_len : g_planesize;
        Obj *p = new_obj(ak, len, es * len);
        memcpy(p->data, g_planes + ((size_t)g_next * g_nviews + v) * g_planesize * es, es * len);
        ((Obj **)arr->data)[v] = p;
    }
    g_next++;
    return arr;
}
static void CallStaticVoidMethod(void *env, void *cls, void *mid, ...)
{
    (void)env; (void)cls;
    va_list ap;
    va_start(ap, mid);
    Obj *o = va_arg(ap, Obj *);
    va_end(ap);
    if (mid != &MID_RET || !o) return;
    if (g_pending) g_unexpected++; /* JNI forbids calls with an exception pending */
    log_call('R');
    if (g_nret == g_ret_cap) {
        g_ret_cap = g_ret_cap ? 2 * g_ret_cap : 16;
        g_ret = (unsigned char **)realloc(g_ret, sizeof(unsigned char *) * g_ret_cap);
    }
    size_t bytes = elem_size(o->kind) * (size_t)o->len;
    g_ret[g_nret] = (unsigned char *)malloc(bytes ? bytes : 1);
    memcpy(g_ret[g_nret], o->data, bytes);
    g_nret++;
}
static void *NewStringUTF(void *env, const char *s)
{
    (void)env;
    Obj *o = new_obj(K_STRING, (int)strlen(s), strlen(s) + 1);
    memcpy(o->data, s, strlen(s) + 1);
    return o;
}
static const char *GetStringUTFChars(void *env, Obj *s, uint8_t *copy)
{
    (void)env;
    if (copy) *copy = 0;
    return s ? (const char *)s->data : NULL;
}
static void ReleaseStringUTFChars(void *env, void *s, const char *c) { (void)env; (void)s; (void)c; }
static int GetArrayLength(void *env, Obj *a) { (void)env; return a ? a->len : 0; }
static void *GetObjectArrayElement(void *env, Obj *a, int i)
{
    (void)env;
    if (!a || a->kind != K_OBJARR || i < 0 || i >= a->len) { g_unexpected++; return NULL; }
    return ((Obj **)a->data)[i];
}
static void *NewByteArray(void *env, int n) { (void)env; return new_obj(K_BYTEARR, n, (size_t)n); }
static void *NewShortArray(void *env, int n) { (void)env; return new_obj(K_SHORTARR, n, 2 * (size_t)n); }
static void *GetArrayElements(void *env, Obj *a, uint8_t *copy)
{
    (void)env;
    if (copy) *copy = 0;
    return a ? a->data : NULL;
}
static void ReleaseArrayElements(void *env, void *a, void *p, int mode) { (void)env; (void)a; (void)p; (void)mode; }
static void SetArrayRegion(void *env, Obj *a, int start, int len, const void *buf)
{
    (void)env;
    size_t es = elem_size(a->kind);
    memcpy((char *)a->data + es * start, buf, es * len);
}
static int GetJavaVM(void *env, void ***vm) { (void)env; *vm = (void **)&g_vm; return 0; }
static void *GetPrimitiveArrayCritical(void *env, Obj *a, uint8_t *copy)
{
    (void)env;
    if (copy) *copy = 0;
    return a ? a->data : NULL;
}
static void ReleasePrimitiveArrayCritical(void *env, void *a, void *p, int mode) { (void)env; (void)a; (void)p; (void)mode; }

/* ---- JavaVM ---- */
static int AttachCurrentThread(void *vm, void **penv, void *args)
{
    (void)vm; (void)args;
    *penv = (void *)&g_env;
    return 0;
}

static void init_tables(void)
{
    if (g_env) return;
    for (int i = 0; i < 240; ++i) g_table[i] = (void *)unexpected;
    g_table[6] = (void *)FindClass;
    g_table[14] = (void *)ThrowNew;
    g_table[21] = (void *)NewGlobalRef;
    g_table[22] = (void *)DeleteGlobalRef;
    g_table[23] = (void *)DeleteLocalRef;
    g_table[113] = (void *)GetStaticMethodID;
    g_table[114] = (void *)CallStaticObjectMethod;
    g_table[141] = (void *)CallStaticVoidMethod;
    g_table[167] = (void *)NewStringUTF;
    g_table[169] = (void *)GetStringUTFChars;
    g_table[170] = (void *)ReleaseStringUTFChars;
    g_table[171] = (void *)GetArrayLength;
    g_table[173] = (void *)GetObjectArrayElement;
    g_table[176] = (void *)NewByteArray;
    g_table[178] = (void *)NewShortArray;
    g_table[184] = (void *)GetArrayElements;  /* byte */
    g_table[186] = (void *)GetArrayElements;  /* short */
    g_table[189] = (void *)GetArrayElements;  /* float */
    g_table[192] = (void *)ReleaseArrayElements;
    g_table[194] = (void *)ReleaseArrayElements;
    g_table[197] = (void *)ReleaseArrayElements;
    g_table[208] = (void *)SetArrayRegion;    /* byte */
    g_table[210] = (void *)SetArrayRegion;    /* short */
    g_table[219] = (void *)GetJavaVM;
    g_table[222] = (void *)GetPrimitiveArrayCritical;
    g_table[223] = (void *)ReleasePrimitiveArrayCritical;
    g_table[228] = (void *)ExceptionCheck;
    g_env = (void *const *)g_table;
    for (int i = 0; i < 8; ++i) g_vm_table[i] = (void *)unexpected;
    g_vm_table[4] = (void *)AttachCurrentThread;
    g_vm = (void *const *)g_vm_table;
}

/* ---- harness API (called from the tests through ctypes) ---- */
void *fj_env(void) { init_tables(); return (void *)&g_env; }
void *fj_class(void) { return &g_class; }
void *fj_string(const char *s) { init_tables(); return NewStringUTF(NULL, s); }
void *fj_object_array(int n) { return new_obj(K_OBJARR, n, sizeof(Obj *) * (size_t)n); }
void fj_object_array_set(Obj *a, int i, Obj *e) { ((Obj **)a->data)[i] = e; }
void *fj_short_array(int n, const int16_t *d)
{
    Obj *o = new_obj(K_SHORTARR, n, 2 * (size_t)n);
    if (d) memcpy(o->data, d, 2 * (size_t)n);
    return o;
}
void *fj_byte_array(int n, const int8_t *d)
{
    Obj *o = new_obj(K_BYTEARR, n, (size_t)n);
    if (d) memcpy(o->data, d, (size_t)n);
    return o;
}
void *fj_float_array(int n, const float *d)
{
    Obj *o = new_obj(K_FLOATARR, n, 4 * (size_t)n);
    if (d) memcpy(o->data, d, 4 * (size_t)n);
    return o;
}
const char *fj_pending_exception(void) { return g_pending ? g_exception : NULL; }
void fj_clear_exception(void) { g_pending = 0; g_exception[0] = 0; }
int fj_unexpected_calls(void) { return g_unexpected; }
int fj_local_refs_deleted(void) { return g_local_deleted; }
int fj_global_refs(void) { return g_global_refs; }
void fj_set_methods_present(int on) { g_methods_present = on; }

void fj_set_planes(int kind, int nplanes, int nviews, int planesize, const void *data)
{
    free(g_planes);
    size_t es = kind == 16 ? 2 : 1;
    size_t bytes = es * (size_t)nplanes * nviews * planesize;
    g_planes = (unsigned char *)malloc(bytes ? bytes : 1);
    if (bytes) memcpy(g_planes, data, bytes);
    g_kind = kind;
    g_nplanes = nplanes;
    g_nviews = nviews;
    g_planesize = planesize;
    g_next = 0;
}
int fj_planes_served(void) { return g_next; }
int fj_returned_count(void) { return g_nret; }
const void *fj_returned_plane(int i) { return (i >= 0 && i < g_nret) ? g_ret[i] : NULL; }

const char *fj_call_log(void) { return g_log; }
void fj_set_short_view(int plane, int view, int len)
{
    g_short_plane = plane;
    g_short_view = view;
    g_short_len = len;
}

void fj_reset(void)
{
    g_loglen = 0;
    g_log[0] = 0;
    g_short_plane = -1;
    while (g_objs) {
        Obj *n = g_objs->next;
        free(g_objs->data);
        free(g_objs);
        g_objs = n;
    }
    for (int i = 0; i < g_nret; ++i) free(g_ret[i]);
    g_nret = 0;
    g_pending = 0;
    g_exception[0] = 0;
    g_unexpected = 0;
    g_local_deleted = 0;
    g_methods_present = 1;
    g_next = 0;
}
const char *fj_string_value(Obj *s) { return (s && s->kind == K_STRING) ? (const char *)s->data : NULL; }
