// JNI binding for the Fiji plugin class fastspim.NativeSPIMReconstructionCuda.
// Reference: fastspim_NativeSPIMReconstructionCuda.cpp:13-538 and the
// javah header fastspim_NativeSPIMReconstructionCuda.h:11-110.
//
// Same entry points, signatures and exception messages as the reference.
// Differences (all on error paths or JNI hygiene): bit depth is validated
// before any work; missing static callback methods raise instead of leaving
// a half-initialised session; the calling class is kept as a global
// reference (a local reference does not survive the native call); local
// references created per plane are deleted; the interactive call raises if
// no session exists; failed file opens raise instead of crashing.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fastspim_NativeSPIMReconstructionCuda.h"
#include "../../include/fmvd_cuda_utils.h"
#include "../../include/fmvd_deconvolve.h"
#include "../../include/fmvd_transform.h"

namespace {

template <typename F> F slot(JNIEnv* env, int i) { return reinterpret_cast<F>((*env)[i]); }

jclass FindClass(JNIEnv* env, const char* n) { return slot<jclass (*)(JNIEnv*, const char*)>(env, JNI_FindClass)(env, n); }
jint ThrowNew(JNIEnv* env, jclass c, const char* m)
{
    return slot<jint (*)(JNIEnv*, jclass, const char*)>(env, JNI_ThrowNew)(env, c, m);
}
jboolean ExceptionCheck(JNIEnv* env) { return slot<jboolean (*)(JNIEnv*)>(env, JNI_ExceptionCheck)(env); }
jobject NewGlobalRef(JNIEnv* env, jobject o) { return slot<jobject (*)(JNIEnv*, jobject)>(env, JNI_NewGlobalRef)(env, o); }
void DeleteGlobalRef(JNIEnv* env, jobject o) { slot<void (*)(JNIEnv*, jobject)>(env, JNI_DeleteGlobalRef)(env, o); }
void DeleteLocalRef(JNIEnv* env, jobject o)
{
    if (o) slot<void (*)(JNIEnv*, jobject)>(env, JNI_DeleteLocalRef)(env, o);
}
jmethodID GetStaticMethodID(JNIEnv* env, jclass c, const char* n, const char* s)
{
    return slot<jmethodID (*)(JNIEnv*, jclass, const char*, const char*)>(env, JNI_GetStaticMethodID)(env, c, n, s);
}
jobject CallStaticObjectMethod0(JNIEnv* env, jclass c, jmethodID m)
{
    return slot<jobject (*)(JNIEnv*, jclass, jmethodID, ...)>(env, JNI_CallStaticObjectMethod)(env, c, m);
}
void CallStaticVoidMethod1(JNIEnv* env, jclass c, jmethodID m, jobject a)
{
    slot<void (*)(JNIEnv*, jclass, jmethodID, ...)>(env, JNI_CallStaticVoidMethod)(env, c, m, a);
}
jstring NewStringUTF(JNIEnv* env, const char* s) { return slot<jstring (*)(JNIEnv*, const char*)>(env, JNI_NewStringUTF)(env, s); }
const char* GetStringUTFChars(JNIEnv* env, jstring s)
{
    return slot<const char* (*)(JNIEnv*, jstring, jboolean*)>(env, JNI_GetStringUTFChars)(env, s, nullptr);
}
void ReleaseStringUTFChars(JNIEnv* env, jstring s, const char* c)
{
    slot<void (*)(JNIEnv*, jstring, const char*)>(env, JNI_ReleaseStringUTFChars)(env, s, c);
}
jsize GetArrayLength(JNIEnv* env, jarray a) { return slot<jsize (*)(JNIEnv*, jarray)>(env, JNI_GetArrayLength)(env, a); }
jobject GetObjectArrayElement(JNIEnv* env, jobjectArray a, jsize i)
{
    return slot<jobject (*)(JNIEnv*, jobjectArray, jsize)>(env, JNI_GetObjectArrayElement)(env, a, i);
}
jshortArray NewShortArray(JNIEnv* env, jsize n) { return slot<jshortArray (*)(JNIEnv*, jsize)>(env, JNI_NewShortArray)(env, n); }
jbyteArray NewByteArray(JNIEnv* env, jsize n) { return slot<jbyteArray (*)(JNIEnv*, jsize)>(env, JNI_NewByteArray)(env, n); }
jshort* GetShortArrayElements(JNIEnv* env, jshortArray a)
{
    return slot<jshort* (*)(JNIEnv*, jshortArray, jboolean*)>(env, JNI_GetShortArrayElements)(env, a, nullptr);
}
jbyte* GetByteArrayElements(JNIEnv* env, jbyteArray a)
{
    return slot<jbyte* (*)(JNIEnv*, jbyteArray, jboolean*)>(env, JNI_GetByteArrayElements)(env, a, nullptr);
}
jfloat* GetFloatArrayElements(JNIEnv* env, jfloatArray a)
{
    return slot<jfloat* (*)(JNIEnv*, jfloatArray, jboolean*)>(env, JNI_GetFloatArrayElements)(env, a, nullptr);
}
void ReleaseShortArrayElements(JNIEnv* env, jshortArray a, jshort* p, jint mode)
{
    slot<void (*)(JNIEnv*, jshortArray, jshort*, jint)>(env, JNI_ReleaseShortArrayElements)(env, a, p, mode);
}
void ReleaseByteArrayElements(JNIEnv* env, jbyteArray a, jbyte* p, jint mode)
{
    slot<void (*)(JNIEnv*, jbyteArray, jbyte*, jint)>(env, JNI_ReleaseByteArrayElements)(env, a, p, mode);
}
void ReleaseFloatArrayElements(JNIEnv* env, jfloatArray a, jfloat* p, jint mode)
{
    slot<void (*)(JNIEnv*, jfloatArray, jfloat*, jint)>(env, JNI_ReleaseFloatArrayElements)(env, a, p, mode);
}
void SetShortArrayRegion(JNIEnv* env, jshortArray a, jsize s, jsize n, const jshort* b)
{
    slot<void (*)(JNIEnv*, jshortArray, jsize, jsize, const jshort*)>(env, JNI_SetShortArrayRegion)(env, a, s, n, b);
}
void SetByteArrayRegion(JNIEnv* env, jbyteArray a, jsize s, jsize n, const jbyte* b)
{
    slot<void (*)(JNIEnv*, jbyteArray, jsize, jsize, const jbyte*)>(env, JNI_SetByteArrayRegion)(env, a, s, n, b);
}
jint GetJavaVM(JNIEnv* env, JavaVM** vm) { return slot<jint (*)(JNIEnv*, JavaVM**)>(env, JNI_GetJavaVM)(env, vm); }
void* GetPrimitiveArrayCritical(JNIEnv* env, jarray a)
{
    return slot<void* (*)(JNIEnv*, jarray, jboolean*)>(env, JNI_GetPrimitiveArrayCritical)(env, a, nullptr);
}
void ReleasePrimitiveArrayCritical(JNIEnv* env, jarray a, void* p, jint mode)
{
    slot<void (*)(JNIEnv*, jarray, void*, jint)>(env, JNI_ReleasePrimitiveArrayCritical)(env, a, p, mode);
}
JNIEnv* AttachCurrentThread(JavaVM* vm)
{
    JNIEnv* env = nullptr;
    reinterpret_cast<jint (*)(JavaVM*, void**, void*)>((*vm)[JNI_VM_AttachCurrentThread])(vm, (void**)&env, nullptr);
    return env;
}

// fastspim_NativeSPIMReconstructionCuda.cpp:13-24
void ThrowException(void* env_ptr, const char* message)
{
    JNIEnv* env = (JNIEnv*)env_ptr;
    jclass cl = FindClass(env, "java/lang/RuntimeException");
    ThrowNew(env, cl, message);
}

void setJavaExceptionHandler(JNIEnv* env) { setErrorHandler(ThrowException, env); }

// Reads `count` floats from each of n files named by a Java String[].
bool read_float_files(JNIEnv* env, jobjectArray names, int n, size_t count, std::vector<std::vector<float>>& out)
{
    if (n < 1 || names == nullptr || GetArrayLength(env, names) < n) {
        ThrowException(env, "file name array holds fewer than nViews entries");
        return false;
    }
    out.assign(n, std::vector<float>(count));
    for (int v = 0; v < n; ++v) {
        jstring js = (jstring)GetObjectArrayElement(env, names, v);
        const char* path = GetStringUTFChars(env, js);
        FILE* f = fopen(path, "rb");
        size_t got = 0;
        if (f) {
            got = fread(out[v].data(), sizeof(float), count, f);
            fclose(f);
        }
        std::string p = path ? path : "";
        ReleaseStringUTFChars(env, js, path);
        DeleteLocalRef(env, js);
        if (!f) {
            ThrowException(env, ("Could not open " + p + " for reading").c_str());
            return false;
        }
        if (got != count) {
            ThrowException(env, ("Short read from " + p).c_str());
            return false;
        }
    }
    return true;
}

// ------------------------------------------------------- interactive state
// fastspim_NativeSPIMReconstructionCuda.cpp:257-321: one session per depth.
struct Interactive {
    void* plan = nullptr;
    JavaVM* jvm = nullptr;
    jmethodID getNextPlane = nullptr, returnNextPlane = nullptr;
    jclass callingClass = nullptr;  // global reference
    int planesize = 0;
    int nViews = 0;
};
Interactive* ia_8 = nullptr;
Interactive* ia_16 = nullptr;

// fastspim_NativeSPIMReconstructionCuda.cpp:323-377
template <typename S> int get_next_plane(Interactive* ia, S** data, int offset)
{
    JNIEnv* env = AttachCurrentThread(ia->jvm);
    jobjectArray jarr = (jobjectArray)CallStaticObjectMethod0(env, ia->callingClass, ia->getNextPlane);
    if (jarr == nullptr || ExceptionCheck(env)) {
        DeleteLocalRef(env, jarr);
        return 0;
    }
    // the upcall's array must hold nViews planes of planesize samples; a
    // short one ends the stream with a Java exception instead of reading
    // past it (the reference copied without checking)
    if (GetArrayLength(env, jarr) < ia->nViews) {
        DeleteLocalRef(env, jarr);
        ThrowException(env, "getNextPlane returned fewer planes than views");
        return 0;
    }
    for (int v = 0; v < ia->nViews; ++v) {
        jarray view = (jarray)GetObjectArrayElement(env, jarr, v);
        if (view == nullptr || GetArrayLength(env, view) < ia->planesize) {
            DeleteLocalRef(env, view);
            DeleteLocalRef(env, jarr);
            ThrowException(env, "getNextPlane returned a view with fewer than planesize samples");
            return 0;
        }
        void* src = GetPrimitiveArrayCritical(env, view);
        if (!src) {
            DeleteLocalRef(env, view);
            DeleteLocalRef(env, jarr);
            ThrowException(env, "not enough memory");
            return 0;
        }
        memcpy(data[v] + offset, src, (size_t)ia->planesize * sizeof(S));
        ReleasePrimitiveArrayCritical(env, view, src, JNI_ABORT);
        DeleteLocalRef(env, view);
    }
    DeleteLocalRef(env, jarr);
    return 1;
}

int get_next_plane_8(unsigned char** d, int o) { return get_next_plane<unsigned char>(ia_8, d, o); }
int get_next_plane_16(unsigned short** d, int o) { return get_next_plane<unsigned short>(ia_16, d, o); }

// With a Java exception pending (a failed getNextPlane) no further upcall
// may be made; the remaining planes are dropped and the exception reaches
// the caller when the native method returns.
void return_next_plane_8(unsigned char* data)
{
    JNIEnv* env = AttachCurrentThread(ia_8->jvm);
    if (ExceptionCheck(env)) return;
    jbyteArray a = NewByteArray(env, ia_8->planesize);
    SetByteArrayRegion(env, a, 0, ia_8->planesize, (const jbyte*)data);
    CallStaticVoidMethod1(env, ia_8->callingClass, ia_8->returnNextPlane, a);
    DeleteLocalRef(env, a);
}

void return_next_plane_16(unsigned short* data)
{
    JNIEnv* env = AttachCurrentThread(ia_16->jvm);
    if (ExceptionCheck(env)) return;
    jshortArray a = NewShortArray(env, ia_16->planesize);
    SetShortArrayRegion(env, a, 0, ia_16->planesize, (const jshort*)data);
    CallStaticVoidMethod1(env, ia_16->callingClass, ia_16->returnNextPlane, a);
    DeleteLocalRef(env, a);
}

template <typename S, typename JArr, typename GetF, typename RelF, typename TransformF>
void transform_impl(JNIEnv* env, jobjectArray data, jint w, jint h, jint d, jfloatArray invMatrix,
                    jint tw, jint th, jint td, jstring outfile, jboolean createMasks, jint border,
                    jfloat zspacing, jstring maskfile, GetF get, RelF rel, TransformF run)
{
    setJavaExceptionHandler(env);
    if (d < 1 || w < 1 || h < 1) {
        ThrowException(env, "invalid source dimensions");
        return;
    }
    if (data == nullptr || GetArrayLength(env, data) < d) {
        ThrowException(env, "data holds fewer than d planes");
        return;
    }
    if (invMatrix == nullptr || GetArrayLength(env, invMatrix) < 12) {
        ThrowException(env, "invMatrix must hold 12 values");
        return;
    }
    std::vector<JArr> jdata(d);
    std::vector<S*> cdata(d);
    for (int z = 0; z < d; ++z) {
        jdata[z] = (JArr)GetObjectArrayElement(env, data, z);
        if (jdata[z] == nullptr || GetArrayLength(env, jdata[z]) < (long long)w * h) {
            for (int k = 0; k < z; ++k) rel(env, jdata[k], cdata[k], JNI_ABORT);
            ThrowException(env, "a plane holds fewer than w*h samples");
            return;
        }
        cdata[z] = (S*)get(env, jdata[z]);
        if (!cdata[z]) {
            for (int k = 0; k < z; ++k) rel(env, jdata[k], cdata[k], JNI_ABORT);
            ThrowException(env, "not enough memory");
            return;
        }
    }
    jfloat* mat = GetFloatArrayElements(env, invMatrix);
    const char* outpath = GetStringUTFChars(env, outfile);
    const char* maskpath = createMasks ? GetStringUTFChars(env, maskfile) : nullptr;
    run(cdata.data(), w, h, d, tw, th, td, mat, outpath, createMasks ? 1 : 0, border, zspacing, maskpath);
    for (int z = 0; z < d; ++z) {
        rel(env, jdata[z], cdata[z], JNI_ABORT);
        DeleteLocalRef(env, jdata[z]);
    }
    ReleaseFloatArrayElements(env, invMatrix, mat, JNI_ABORT);
    ReleaseStringUTFChars(env, outfile, outpath);
    if (maskpath) ReleaseStringUTFChars(env, maskfile, maskpath);
}

}  // namespace

extern "C" {

JNIEXPORT jint JNICALL Java_fastspim_NativeSPIMReconstructionCuda_getNumCudaDevices(JNIEnv* env, jclass)
{
    setJavaExceptionHandler(env);
    return getNumCUDADevices();
}

JNIEXPORT jstring JNICALL Java_fastspim_NativeSPIMReconstructionCuda_getCudaDeviceName(JNIEnv* env, jclass,
                                                                                      jint deviceIdx)
{
    char name[256] = {0};
    setJavaExceptionHandler(env);
    getCudaDeviceName(deviceIdx, name);
    return NewStringUTF(env, name);
}

JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_setCudaDevice(JNIEnv* env, jclass, jint deviceIdx)
{
    setJavaExceptionHandler(env);
    setCudaDevice(deviceIdx);
}

JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_transform16(
    JNIEnv* env, jclass, jobjectArray data, jint w, jint h, jint d, jfloatArray invMatrix, jint targetW,
    jint targetH, jint targetD, jstring outfile, jboolean createTransformedMasks, jint border,
    jfloat zspacing, jstring maskfile)
{
    transform_impl<unsigned short, jshortArray>(
        env, data, w, h, d, invMatrix, targetW, targetH, targetD, outfile, createTransformedMasks, border,
        zspacing, maskfile, [](JNIEnv* e, jshortArray a) { return GetShortArrayElements(e, a); },
        [](JNIEnv* e, jshortArray a, unsigned short* p, jint m) { ReleaseShortArrayElements(e, a, (jshort*)p, m); },
        transform_cuda_16);
}

JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_transform8(
    JNIEnv* env, jclass, jobjectArray data, jint w, jint h, jint d, jfloatArray invMatrix, jint targetW,
    jint targetH, jint targetD, jstring outfile, jboolean createTransformedMasks, jint border,
    jfloat zspacing, jstring maskfile)
{
    transform_impl<unsigned char, jbyteArray>(
        env, data, w, h, d, invMatrix, targetW, targetH, targetD, outfile, createTransformedMasks, border,
        zspacing, maskfile, [](JNIEnv* e, jbyteArray a) { return GetByteArrayElements(e, a); },
        [](JNIEnv* e, jbyteArray a, unsigned char* p, jint m) { ReleaseByteArrayElements(e, a, (jbyte*)p, m); },
        transform_cuda_8);
}

// fastspim_NativeSPIMReconstructionCuda.cpp:171-255
JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_deconvolve(
    JNIEnv* env, jclass, jobjectArray inputfiles, jstring outputfile, jint dataW, jint dataH, jint dataD,
    jobjectArray weightfiles, jobjectArray kernelfiles, jint kernelH, jint kernelW, jint iterationType,
    jint nViews, jint iterations, jint bitDepth)
{
    setJavaExceptionHandler(env);
    if (bitDepth != 8 && bitDepth != 16) {
        ThrowException(env, "Unsupported bit depth");  // :237
        return;
    }
    setCudaDevice(0);  // :189 (the reference always uses device 0 here)
    if (ExceptionCheck(env)) return;
    std::vector<std::vector<float>> kernels, weights;
    if (!read_float_files(env, kernelfiles, nViews, (size_t)kernelW * kernelH, kernels)) return;
    if (!read_float_files(env, weightfiles, nViews, (size_t)dataW * dataH, weights)) return;
    std::vector<float*> kp(nViews), wp(nViews);
    for (int v = 0; v < nViews; ++v) {
        kp[v] = kernels[v].data();
        wp[v] = weights[v].data();
    }
    if (inputfiles == nullptr || GetArrayLength(env, inputfiles) < nViews) {
        ThrowException(env, "file name array holds fewer than nViews entries");
        return;
    }
    std::vector<FILE*> files(nViews, nullptr);
    auto close_all = [&] {
        for (FILE* f : files)
            if (f) fclose(f);
    };
    for (int v = 0; v < nViews; ++v) {
        jstring js = (jstring)GetObjectArrayElement(env, inputfiles, v);
        const char* path = GetStringUTFChars(env, js);
        files[v] = fopen(path, "rb");
        std::string p = path ? path : "";
        ReleaseStringUTFChars(env, js, path);
        DeleteLocalRef(env, js);
        if (!files[v]) {
            close_all();
            ThrowException(env, ("Could not open " + p + " for reading").c_str());
            return;
        }
    }
    const char* opath = GetStringUTFChars(env, outputfile);
    FILE* result = fopen(opath, "wb");
    std::string op = opath ? opath : "";
    ReleaseStringUTFChars(env, outputfile, opath);
    if (!result) {
        close_all();
        ThrowException(env, ("Could not open " + op + " for writing").c_str());
        return;
    }
    const fmvd_psf_type t = (fmvd_psf_type)iterationType;
    if (bitDepth == 8)
        fmvd_deconvolve_files_cuda_8(files.data(), result, dataW, dataH, dataD, wp.data(), kp.data(), kernelH,
                                     kernelW, t, nViews, iterations);
    else
        fmvd_deconvolve_files_cuda_16(files.data(), result, dataW, dataH, dataD, wp.data(), kp.data(), kernelH,
                                      kernelW, t, nViews, iterations);
    close_all();
    fclose(result);
}

// fastspim_NativeSPIMReconstructionCuda.cpp:379-403
JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_deconvolve_1quit8(JNIEnv* env, jclass)
{
    if (ia_8 == nullptr) {
        ThrowException(env, "deconvolution already quit");
        return;
    }
    fmvd_destroy_cuda_8((fmvd_plan_cuda_8*)ia_8->plan);
    DeleteGlobalRef(env, ia_8->callingClass);
    delete ia_8;
    ia_8 = nullptr;
}

JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_deconvolve_1quit16(JNIEnv* env, jclass)
{
    if (ia_16 == nullptr) {
        ThrowException(env, "deconvolution already quit");
        return;
    }
    fmvd_destroy_cuda_16((fmvd_plan_cuda_16*)ia_16->plan);
    DeleteGlobalRef(env, ia_16->callingClass);
    delete ia_16;
    ia_16 = nullptr;
}

// fastspim_NativeSPIMReconstructionCuda.cpp:405-419
JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_deconvolve_1interactive8(JNIEnv* env, jclass,
                                                                                          jint iterations)
{
    if (ia_8 == nullptr) {
        ThrowException(env, "deconvolution not initialized");
        return;
    }
    setJavaExceptionHandler(env);
    fmvd_deconvolve_planes_cuda_8((fmvd_plan_cuda_8*)ia_8->plan, iterations);
}

JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_deconvolve_1interactive16(JNIEnv* env, jclass,
                                                                                           jint iterations)
{
    if (ia_16 == nullptr) {
        ThrowException(env, "deconvolution not initialized");
        return;
    }
    setJavaExceptionHandler(env);
    fmvd_deconvolve_planes_cuda_16((fmvd_plan_cuda_16*)ia_16->plan, iterations);
}

// fastspim_NativeSPIMReconstructionCuda.cpp:421-538
JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_deconvolve_1init(
    JNIEnv* env, jclass clazz, jint dataW, jint dataH, jint dataD, jobjectArray weightfiles,
    jobjectArray kernelfiles, jint kernelH, jint kernelW, jint iterationType, jint nViews, jint bitDepth)
{
    if (bitDepth != 8 && bitDepth != 16) {
        ThrowException(env, "Unsupported bit depth");
        return;
    }
    if ((bitDepth == 8 && ia_8 != nullptr) || (bitDepth == 16 && ia_16 != nullptr)) {
        ThrowException(env, "deconvolution already initialized");  // :438,:444
        return;
    }
    std::vector<std::vector<float>> kernels, weights;
    if (!read_float_files(env, kernelfiles, nViews, (size_t)kernelW * kernelH, kernels)) return;
    if (!read_float_files(env, weightfiles, nViews, (size_t)dataW * dataH, weights)) return;
    jmethodID get = GetStaticMethodID(env, clazz, "getNextPlane", "()[Ljava/lang/Object;");
    jmethodID ret = GetStaticMethodID(env, clazz, "returnNextPlane", "(Ljava/lang/Object;)V");
    if (get == nullptr || ret == nullptr) {
        if (!ExceptionCheck(env)) ThrowException(env, "error in callback: Method not found");
        return;
    }
    JavaVM* jvm = nullptr;
    GetJavaVM(env, &jvm);
    std::vector<float*> kp(nViews), wp(nViews);
    for (int v = 0; v < nViews; ++v) {
        kp[v] = kernels[v].data();
        wp[v] = weights[v].data();
    }
    setJavaExceptionHandler(env);
    const int nStreams = dataD < 3 ? (dataD < 1 ? 1 : dataD) : 3;  // :482
    Interactive* ia = new Interactive();
    ia->jvm = jvm;
    ia->getNextPlane = get;
    ia->returnNextPlane = ret;
    ia->planesize = dataW * dataH;
    ia->nViews = nViews;
    const fmvd_psf_type t = (fmvd_psf_type)iterationType;
    if (bitDepth == 8)
        ia->plan = fmvd_initialize_cuda_8(dataH, dataW, wp.data(), kp.data(), kernelH, kernelW, t, nViews,
                                          nStreams, get_next_plane_8, return_next_plane_8);
    else
        ia->plan = fmvd_initialize_cuda_16(dataH, dataW, wp.data(), kp.data(), kernelH, kernelW, t, nViews,
                                           nStreams, get_next_plane_16, return_next_plane_16);
    if (!ia->plan) {  // already reported as a RuntimeException
        delete ia;
        return;
    }
    // The Java side may be a bounded producer: keep the reference's call
    // order, at most nStreams planes pulled ahead of returnNextPlane
    // (fastspim_NativeSPIMReconstructionCuda.cpp:482, fmvd_deconvolve.impl.cpp:216-245).
    // FMVD_JNI_LOOKAHEAD overrides the bound; 0 selects the batched pipeline.
    int lookahead = nStreams;
    if (const char* e = getenv("FMVD_JNI_LOOKAHEAD"); e && *e) lookahead = atoi(e);
    if (bitDepth == 8) fmvd_set_lookahead_cuda_8((fmvd_plan_cuda_8*)ia->plan, lookahead);
    else fmvd_set_lookahead_cuda_16((fmvd_plan_cuda_16*)ia->plan, lookahead);
    ia->callingClass = (jclass)NewGlobalRef(env, clazz);
    if (bitDepth == 8) ia_8 = ia;
    else ia_16 = ia;
}

}  // extern "C"
