/*
 * JNI entry points of libNativeSPIMReconstructionCuda.so for the Java class
 * fastspim.NativeSPIMReconstructionCuda (all static native methods).
 * Drop-in for fastspim_NativeSPIMReconstructionCuda.h:11-110 of
 * bene51/gpu_deconvolution: same symbols and Java signatures.  Declared over
 * the ABI in fmvd_jni_abi.h because no JDK header is available here; when a
 * JDK's <jni.h> is used instead the types and layouts are identical.
 */
#ifndef FASTSPIM_NATIVESPIMRECONSTRUCTIONCUDA_H
#define FASTSPIM_NATIVESPIMRECONSTRUCTIONCUDA_H

#include "fmvd_jni_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* int getNumCudaDevices()  (reference .h:11-13, .cpp:26-32) */
JNIEXPORT jint JNICALL Java_fastspim_NativeSPIMReconstructionCuda_getNumCudaDevices(JNIEnv *env, jclass);

/* String getCudaDeviceName(int)  (.h:15-18, .cpp:34-44) */
JNIEXPORT jstring JNICALL Java_fastspim_NativeSPIMReconstructionCuda_getCudaDeviceName(JNIEnv *env, jclass,
                                                                                      jint deviceIdx);

/* void setCudaDevice(int)  (.h:20-23, .cpp:46-53) */
JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_setCudaDevice(JNIEnv *env, jclass,
                                                                               jint deviceIdx);

/* void transform8/16(byte[][]/short[][] data, int w, int h, int d, float[] invMatrix,
 *   int targetW, int targetH, int targetD, String outfile, boolean createTransformedMasks,
 *   int border, float zspacing, String maskfile)  (.h:25-57, .cpp:55-169) */
JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_transform8(
    JNIEnv *env, jclass, jobjectArray data, jint w, jint h, jint d, jfloatArray invMatrix, jint targetW,
    jint targetH, jint targetD, jstring outfile, jboolean createTransformedMasks, jint border,
    jfloat zspacing, jstring maskfile);
JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_transform16(
    JNIEnv *env, jclass, jobjectArray data, jint w, jint h, jint d, jfloatArray invMatrix, jint targetW,
    jint targetH, jint targetD, jstring outfile, jboolean createTransformedMasks, jint border,
    jfloat zspacing, jstring maskfile);

/* void deconvolve(String[] inputfiles, String outputfile, int dataW, int dataH, int dataD,
 *   String[] weightfiles, String[] kernelfiles, int kernelH, int kernelW, int psfType,
 *   int nViews, int iterations, int bitDepth)  (.h:61-76, .cpp:171-255) */
JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_deconvolve(
    JNIEnv *env, jclass, jobjectArray inputfiles, jstring outputfile, jint dataW, jint dataH, jint dataD,
    jobjectArray weightfiles, jobjectArray kernelfiles, jint kernelH, jint kernelW, jint psfType,
    jint nViews, jint iterations, jint bitDepth);

/* void deconvolve_quit8/16()  (.h:78-84, .cpp:379-403) */
JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_deconvolve_1quit8(JNIEnv *env, jclass clazz);
JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_deconvolve_1quit16(JNIEnv *env, jclass clazz);

/* void deconvolve_interactive8/16(int iterations)  (.h:86-94, .cpp:405-419): planes come
 * from the static method Object[] getNextPlane() (null = end of data) and go back through
 * the static method void returnNextPlane(Object) of the calling class. */
JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_deconvolve_1interactive8(JNIEnv *env,
                                                                                          jclass clazz,
                                                                                          jint iterations);
JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_deconvolve_1interactive16(JNIEnv *env,
                                                                                           jclass clazz,
                                                                                           jint iterations);

/* void deconvolve_init(int dataW, int dataH, int dataD, String[] weightfiles,
 *   String[] kernelfiles, int kernelH, int kernelW, int iterationType, int nViews,
 *   int bitDepth)  (.h:98-110, .cpp:421-538) */
JNIEXPORT void JNICALL Java_fastspim_NativeSPIMReconstructionCuda_deconvolve_1init(
    JNIEnv *env, jclass clazz, jint dataW, jint dataH, jint dataD, jobjectArray weightfiles,
    jobjectArray kernelfiles, jint kernelH, jint kernelW, jint iterationType, jint nViews, jint bitDepth);

#ifdef __cplusplus
}
#endif

#endif
