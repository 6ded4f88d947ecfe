// Links the HIP library built by `make -C aether-core-engine_amd`
// (lib/libaether_fluid_hip.so); AETHER_FLUID_HIP_LIB overrides its folder.
fn main() {
    let dir = std::env::var("AETHER_FLUID_HIP_LIB").unwrap_or_else(|_| {
        let manifest = std::env::var("CARGO_MANIFEST_DIR").unwrap_or_else(|_| ".".into());
        format!("{manifest}/../lib")
    });
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=aether_fluid_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
    println!("cargo:rerun-if-env-changed=AETHER_FLUID_HIP_LIB");
}
