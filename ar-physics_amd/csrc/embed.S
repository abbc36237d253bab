/* Embeds the gfx950 code object built from kernels.hip into libarp.so. */
    .section .rodata
    .global arp_kernels_co
    .type arp_kernels_co, @object
    .p2align 12
arp_kernels_co:
    .incbin ARP_CO_PATH
    .byte 0
    .section .note.GNU-stack,"",@progbits
