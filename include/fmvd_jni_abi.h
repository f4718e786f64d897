/*
 * Minimal JNI ABI for the binding in jni_binding.cpp.
 *
 * The image this library is built on has no JDK (no jni.h), so the binding
 * calls the JNIEnv / JavaVM function tables by slot index as the JNI
 * specification ("JNI Functions: Interface Function Table" and "Invocation
 * API: Invocation API Functions") defines them.  A JNIEnv* points at a
 * pointer to the function table (the C layout of jni.h: JNIEnv is
 * `const struct JNINativeInterface_ *`), and a JavaVM* likewise.  The same
 * layout is what tests/jni_fake implements to drive the binding without a JVM.
 */
#ifndef FMVD_JNI_ABI_H
#define FMVD_JNI_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t jint;
typedef int64_t jlong;
typedef int8_t jbyte;
typedef uint8_t jboolean;
typedef uint16_t jchar;
typedef int16_t jshort;
typedef float jfloat;
typedef double jdouble;
typedef jint jsize;

typedef void *jobject;
typedef jobject jclass;
typedef jobject jstring;
typedef jobject jarray;
typedef jobject jobjectArray;
typedef jobject jbyteArray;
typedef jobject jshortArray;
typedef jobject jfloatArray;
typedef jobject jthrowable;
typedef void *jmethodID;

typedef void *const *JNIEnv;  /* *env is the function table */
typedef void *const *JavaVM;  /* *vm is the invoke table */

#define JNI_ABORT 2
#define JNIEXPORT __attribute__((visibility("default")))
#define JNICALL

/* Interface function table slots (JNI specification). */
enum {
    JNI_FindClass = 6,
    JNI_ThrowNew = 14,
    JNI_ExceptionCheck = 228,
    JNI_NewGlobalRef = 21,
    JNI_DeleteGlobalRef = 22,
    JNI_DeleteLocalRef = 23,
    JNI_GetStaticMethodID = 113,
    JNI_CallStaticObjectMethod = 114,
    JNI_CallStaticVoidMethod = 141,
    JNI_NewStringUTF = 167,
    JNI_GetStringUTFChars = 169,
    JNI_ReleaseStringUTFChars = 170,
    JNI_GetArrayLength = 171,
    JNI_GetObjectArrayElement = 173,
    JNI_NewByteArray = 176,
    JNI_NewShortArray = 178,
    JNI_GetByteArrayElements = 184,
    JNI_GetShortArrayElements = 186,
    JNI_GetFloatArrayElements = 189,
    JNI_ReleaseByteArrayElements = 192,
    JNI_ReleaseShortArrayElements = 194,
    JNI_ReleaseFloatArrayElements = 197,
    JNI_SetByteArrayRegion = 208,
    JNI_SetShortArrayRegion = 210,
    JNI_GetJavaVM = 219,
    JNI_GetPrimitiveArrayCritical = 222,
    JNI_ReleasePrimitiveArrayCritical = 223
};

/* Invocation interface slots. */
enum { JNI_VM_AttachCurrentThread = 4, JNI_VM_GetEnv = 6 };

#ifdef __cplusplus
}
#endif

#endif
