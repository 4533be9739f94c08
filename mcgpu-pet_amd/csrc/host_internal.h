// host_internal.h — declarations shared by the host-side translation units.
#ifndef MCPET_HOST_INTERNAL_H
#define MCPET_HOST_INTERNAL_H

const char* mcpet_host_last_error();
void mcpet_host_set_error(const char* msg);

#endif
