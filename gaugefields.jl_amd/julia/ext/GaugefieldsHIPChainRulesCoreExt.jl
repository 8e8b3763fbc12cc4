"""
ChainRules reverse rule of the device stout map: `stout_smear(U, nn, links)` returns the
smeared links of `links` (host arrays); the pullback of a cotangent of them is
gf_stout_backward's chain rule through every exp level (the Λ chain of
stout_fast.jl:317-407), converted between the ChainRules convention (∂/∂Re + i ∂/∂Im) and the
library's dS/dU convention by `GaugefieldsHIP.stout_pullback`.
"""
module GaugefieldsHIPChainRulesCoreExt

using ChainRulesCore
import GaugefieldsHIP: stout_smear, stout_pullback, HIPLinks
import Gaugefields: CovNeuralnet

function ChainRulesCore.rrule(::typeof(stout_smear), U::HIPLinks, nn::CovNeuralnet, links::AbstractVector)
    out = stout_smear(U, nn, links)
    function stout_smear_pullback(Ȳ)
        Ū = unthunk(Ȳ)
        cot = [Ū[μ] isa AbstractZero ? zeros(ComplexF64, size(out[μ])) : Ū[μ] for μ in 1:4]
        return NoTangent(), NoTangent(), NoTangent(), stout_pullback(U, nn, links, cot)
    end
    return out, stout_smear_pullback
end

end # module
