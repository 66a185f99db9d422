/* User message callback.  Interface of reference include/srp/message_callback.h:14-45. */
#pragma once

typedef enum SRPMessageType
{
	SRP_MESSAGE_ERROR,
	SRP_MESSAGE_WARNING
} SRPMessageType;

typedef enum SRPMessageSeverity
{
	SRP_MESSAGE_SEVERITY_LOW,
	SRP_MESSAGE_SEVERITY_MODERATE,
	SRP_MESSAGE_SEVERITY_HIGH
} SRPMessageSeverity;

/** Called by the library on errors and warnings, with the formatted message,
 *  the name of the API function that raised it and the user pointer. */
typedef void (*SRPMessageCallbackFunc)(
	SRPMessageType type, SRPMessageSeverity severity, const char* sourceFunction,
	const char* message, void* userParameter
);

typedef struct SRPMessageCallback
{
	SRPMessageCallbackFunc func;  /**< NULL: messages are dropped */
	void* userParameter;          /**< Passed back to `func` unchanged */
} SRPMessageCallback;
